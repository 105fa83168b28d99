/* Learned warm start for the thermal / electrical solves: a 3-level 3D
 * convolutional encoder-decoder (channels 32/64/128, 3x3x3 kernels, bf16 weights
 * and activations with fp32 accumulation on the gfx950 matrix cores) that maps
 * per-node input fields of a layered stack to a normalised solution guess, and is
 * trained on the GPU by hand-written forward and backward kernels on solutions
 * the solver generates.
 *
 * None of these entry points has a counterpart in the reference (PLaSK starts
 * every solve from the previous solution or from inittemp, therm3d.cpp:79 and
 * iterative_matrix.hpp:205-209); the guess enters a solve through
 * pf_solve / pf_solution_set of plask_fem.h like any other start vector.
 *
 * Grid: nx, ny, nz nodes (multiples of 4, (nx ny nz / 64) a multiple of 32),
 * node index v = (z ny + y) nx + x.  Inputs: 8 channels per node, channels-last
 * ([v][8], fp32 on the host, bf16 on the device).  Output: one fp32 per node.
 * Parameters: one flat fp32 vector (per convolution w[27][cout][cin] then
 * b[cout], in the order C1 8->32, C2 32->64, C3 64->128, C4 128->64, C5 64->32;
 * then the 32 head weights and the head bias).  Status codes as plask_fem.h. */
#ifndef PLASK_WARMNET_H
#define PLASK_WARMNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_NET_OK 0
#define PF_NET_EINVAL (-1)
#define PF_NET_EDEVICE (-3)

typedef struct pf_net pf_net;

/* Network for one grid; max_samples training samples are held on the device. */
int pf_net_create(pf_net** out, int device, int nx, int ny, int nz, int max_samples);
void pf_net_destroy(pf_net* net);
const char* pf_net_last_error(const pf_net* net);
int64_t pf_net_num_params(const pf_net* net);

/* Parameters (fp32 master copy); setting them resets the optimiser state. */
int pf_net_set_params(pf_net* net, const float* params, int64_t n);
int pf_net_get_params(pf_net* net, float* params, int64_t n);

/* Per-node loss weights (0 excludes a node, e.g. Dirichlet nodes). */
int pf_net_set_mask(pf_net* net, const float* mask);

/* Loss, e = y - target, m = mask:
 *   L = [sum m_v e_v^2 + lambda sum_{x,y edges} (e_a - e_b)^2] / sum m_v
 *       + sum_{z edges} w_v (e_v - e_{v + nx ny})^2,
 * with w from pf_net_sample_weights for that sample, else lambda / sum m_v.
 * lambda defaults to 0. */
int pf_net_set_loss(pf_net* net, float lambda);

/* z-edge loss weights of stored sample idx ([v], edge v -- v + nx ny; host fp32). */
int pf_net_sample_weights(pf_net* net, int idx, const float* wz);

/* Training sample idx: inputs [v][8] and target [v] (host fp32). */
int pf_net_sample_set(pf_net* net, int idx, const float* feat, const float* target);

/* Forward pass: inputs from the host, or a stored sample; out = host [v], or NULL to
 * leave the prediction only on the device (pf_net_output_device). */
int pf_net_predict(pf_net* net, const float* feat, float* out);
int pf_net_predict_sample(pf_net* net, int idx, float* out);

/* Device address of the last prediction ([v] fp32, owned by the network, valid until
 * the next forward pass or pf_net_destroy); complete when pf_net_predict returns. */
const float* pf_net_output_device(const pf_net* net);

/* Loss (as pf_net_set_loss) and its gradient with respect to every
 * parameter for stored sample idx (for gradient checks). */
int pf_net_gradient(pf_net* net, int idx, float* grads, double* loss);

/* One Adam step (b1 0.9, b2 0.999, eps 1e-8) on the mean gradient of n stored
 * samples; loss = their mean loss before the step. */
int pf_net_train_batch(pf_net* net, const int* idx, int n, float lr, double* loss);

#ifdef __cplusplus
}
#endif

#endif
