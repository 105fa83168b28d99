// Warm-start network kernels: 3x3x3 convolutions on bf16 MFMA (forward, backward
// data, backward weights), pooling / upsampling, the pointwise output head, the
// loss and Adam.  Tensors are channels-last, voxel v = (z * ny + y) * nx + x,
// element [v][c].  All accumulation is fp32.
//
// MFMA v_mfma_f32_16x16x32_bf16 (gfx950), one wave: D(16x16) += A(16x32) B(32x16).
// Lane l supplies row l%16 of A and column l%16 of B, eight k values each (slot
// j of lane group g = l/16 is k = 8g + j on both operands, so any internal k
// order of the instruction gives the same sum), and receives D rows 4g..4g+3 of
// column l%16.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Vol {
    int nx, ny, nz;
    __host__ __device__ long long n() const { return (long long)nx * ny * nz; }
};

__device__ __forceinline__ float b2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2b(float v) { return (bf16)v; }

// tap t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); tap 26 - t is the mirrored offset
__device__ __forceinline__ void tap_off(int t, int& dx, int& dy, int& dz) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

// Eight consecutive channels of voxel (x, y, z) of the conv input, zero outside
// the volume.  UP: the input is stored at half resolution (nearest upsampling
// fused into the load), src = (x/2, y/2, z/2) of a volume of half the size.
template <bool UP>
__device__ __forceinline__ bf16x8 load_in8(const bf16* __restrict__ in, const Vol& o, int cin, int x, int y, int z,
                                           int c0) {
    bf16x8 v;
    if (x < 0 || y < 0 || z < 0 || x >= o.nx || y >= o.ny || z >= o.nz) {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = f2b(0.f);
        return v;
    }
    long long s;
    if (UP) s = ((long long)(z >> 1) * (o.ny >> 1) + (y >> 1)) * (o.nx >> 1) + (x >> 1);
    else s = ((long long)z * o.ny + y) * o.nx + x;
    return *reinterpret_cast<const bf16x8*>(in + s * cin + c0);
}

// Output-stationary convolution: one wave = 16 consecutive voxels x (NT * 16)
// output channels, K = 27 * CIN in chunks of 32 (CIN % 8 == 0; a chunk may span
// taps when CIN < 32).  w: [27][cout][CIN] bf16.  Epilogue: + bias[co]
// (nullable) + skip[v][co] (nullable), ReLU if RELU, store bf16.
// Launch: 256 threads, one wave per (voxel tile, channel group).
template <int CIN, int NT, bool UP, bool RELU>
__global__ __launch_bounds__(256) void k_conv3(Vol o, const bf16* __restrict__ in, const bf16* __restrict__ w,
                                               const float* __restrict__ bias, const bf16* __restrict__ skip,
                                               bf16* __restrict__ out, int cout) {
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ngroups = cout / (16 * NT);
    const long long ntiles = o.n() / 16;
    if (wave >= ntiles * ngroups) return;
    const long long tile = wave / ngroups;
    const int co0 = (int)(wave % ngroups) * 16 * NT;
    const int r = lane & 15, g = lane >> 4;
    const long long v = tile * 16 + r;  // A row voxel of this lane
    const int x = (int)(v % o.nx), y = (int)((v / o.nx) % o.ny), z = (int)(v / ((long long)o.nx * o.ny));
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int K = 27 * CIN;
    constexpr int NKC = (K + 31) / 32;
    for (int kc = 0; kc < NKC; ++kc) {
        const int kk = kc * 32 + 8 * g;
        const int tap = kk / CIN, c0 = kk - tap * CIN;
        bf16x8 a, b[NT];
        if (tap < 27) {
            int dx, dy, dz;
            tap_off(tap, dx, dy, dz);
            a = load_in8<UP>(in, o, CIN, x + dx, y + dy, z + dz, c0);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                b[t] = *reinterpret_cast<const bf16x8*>(w + ((size_t)tap * cout + co0 + 16 * t + r) * CIN + c0);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = f2b(0.f);
#pragma unroll
            for (int t = 0; t < NT; ++t) b[t] = a;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[t], 0, 0, 0);
    }
    // D rows 4g + i (voxels), column r (channel)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + 16 * t + r;
        const float bb = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long vo = tile * 16 + 4 * g + i;
            float s = acc[t][i] + bb;
            if (skip) s += b2f(skip[vo * cout + co]);
            if (RELU) s = fmaxf(s, 0.f);
            out[vo * cout + co] = f2b(s);
        }
    }
}

// Weight gradient: gw[tap][co][ci] += sum_v gy[v][co] x[v + off(tap)][ci].
// One wave = one tap x 16 output x 16 input channels over a range of voxels
// (K = voxels, 32 per MFMA); fp32 atomics into gw at the end.
// UP: x is stored at half resolution (the fused upsampling of the forward pass).
template <bool UP>
__global__ __launch_bounds__(256) void k_conv3_wgrad(Vol o, const bf16* __restrict__ gy, const bf16* __restrict__ x,
                                                     float* __restrict__ gw, int cout, int cin, int vchunk) {
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nco = cout / 16, nci = (cin + 15) / 16;
    const long long nv = o.n();
    const long long nchunks = (nv + vchunk - 1) / vchunk;
    const long long nw = 27LL * nco * nci * nchunks;
    if (wave >= nw) return;
    long long rem = wave;
    const int chunk = (int)(rem % nchunks);
    rem /= nchunks;
    const int ct = (int)(rem % nci);
    rem /= nci;
    const int ot = (int)(rem % nco);
    const int tap = (int)(rem / nco);
    int dx, dy, dz;
    tap_off(tap, dx, dy, dz);
    const int r = lane & 15, g = lane >> 4;
    const int co = ot * 16 + r, ci = ct * 16 + r;
    const bool ci_ok = ci < cin;
    const long long vb0 = (long long)chunk * vchunk;
    const long long vend = min(nv, vb0 + vchunk);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long vb = vb0; vb < vend; vb += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long long v = vb + 8 * g + j;
            float av = 0.f, bv = 0.f;
            if (v < vend) {
                av = b2f(gy[v * cout + co]);
                const int xx = (int)(v % o.nx) + dx, yy = (int)((v / o.nx) % o.ny) + dy,
                          zz = (int)(v / ((long long)o.nx * o.ny)) + dz;
                if (ci_ok && xx >= 0 && yy >= 0 && zz >= 0 && xx < o.nx && yy < o.ny && zz < o.nz) {
                    long long s;
                    if (UP) s = ((long long)(zz >> 1) * (o.ny >> 1) + (yy >> 1)) * (o.nx >> 1) + (xx >> 1);
                    else s = ((long long)zz * o.ny + yy) * o.nx + xx;
                    bv = b2f(x[s * cin + ci]);
                }
            }
            a[j] = f2b(av);
            b[j] = f2b(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D[m = co row 4g + i][n = ci column r]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int com = ot * 16 + 4 * g + i;
        if (ci_ok) atomicAdd(&gw[((size_t)tap * cout + com) * cin + ci], acc[i]);
    }
}

// Backward data through an upsampled input: gx at half resolution is the sum of
// the gradients of its 8 children.  gfull: [V][c] at full resolution.
__global__ void k_up_bwd(Vol half, const bf16* __restrict__ gfull, bf16* __restrict__ ghalf, int c,
                         const bf16* __restrict__ relu_of) {
    const long long n = half.n() * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % half.nx), y = (int)((v / half.nx) % half.ny), z = (int)(v / ((long long)half.nx * half.ny));
        const int fx = 2 * half.nx, fy = 2 * half.ny;
        float s = 0.f;
        for (int k = 0; k < 8; ++k) {
            const long long f = ((long long)(2 * z + (k >> 2)) * fy + (2 * y + ((k >> 1) & 1))) * fx + 2 * x + (k & 1);
            s += b2f(gfull[f * c + ch]);
        }
        if (relu_of && !(b2f(relu_of[i]) > 0.f)) s = 0.f;
        ghalf[i] = f2b(s);
    }
}

// 2x2x2 average pooling: out (half resolution) = mean of the 8 children.
__global__ void k_pool(Vol half, const bf16* __restrict__ in, bf16* __restrict__ out, int c) {
    const long long n = half.n() * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % half.nx), y = (int)((v / half.nx) % half.ny), z = (int)(v / ((long long)half.nx * half.ny));
        const int fx = 2 * half.nx, fy = 2 * half.ny;
        float s = 0.f;
        for (int k = 0; k < 8; ++k) {
            const long long f = ((long long)(2 * z + (k >> 2)) * fy + (2 * y + ((k >> 1) & 1))) * fx + 2 * x + (k & 1);
            s += b2f(in[f * c + ch]);
        }
        out[i] = f2b(0.125f * s);
    }
}

// Gradient into a ReLU output e that feeds both a skip connection (gradient
// gskip, full resolution) and a 2x2x2 pool (gradient gpool at half resolution):
// gpre = (gskip + gpool[parent] / 8) * (e > 0).
__global__ void k_skip_pool_bwd(Vol full, const bf16* __restrict__ gskip, const bf16* __restrict__ gpool,
                                const bf16* __restrict__ e, bf16* __restrict__ gpre, int c) {
    const long long n = full.n() * c;
    const int hx = full.nx >> 1, hy = full.ny >> 1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % full.nx), y = (int)((v / full.nx) % full.ny), z = (int)(v / ((long long)full.nx * full.ny));
        const long long p = ((long long)(z >> 1) * hy + (y >> 1)) * hx + (x >> 1);
        float s = b2f(gskip[i]) + 0.125f * b2f(gpool[p * c + ch]);
        if (!(b2f(e[i]) > 0.f)) s = 0.f;
        gpre[i] = f2b(s);
    }
}

// Per-channel sum over voxels (bias gradients): gb[ch] += sum_v g[v][ch].
__global__ void k_chan_sum(long long nv, int c, const bf16* __restrict__ g, float* __restrict__ gb) {
    __shared__ float red[256];
    // thread layout: c columns, 256 / c rows of voxels
    const int rows = 256 / c;
    const int ch = threadIdx.x % c, row = threadIdx.x / c;
    float s = 0.f;
    if (row < rows)
        for (long long v = (long long)blockIdx.x * rows + row; v < nv; v += (long long)gridDim.x * rows)
            s += b2f(g[v * c + ch]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < c) {
        float t = 0.f;
        for (int k = 0; k < rows; ++k) t += red[k * c + threadIdx.x];
        atomicAdd(&gb[threadIdx.x], t);
    }
}

// Output head y[v] = sum_c d[v][c] h[c] + hb (32 channels), fp32 out.
__global__ void k_head(long long nv, const bf16* __restrict__ d, const float* __restrict__ h, const float* __restrict__ hb,
                       float* __restrict__ y) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (long long)gridDim.x * blockDim.x) {
        const bf16x8* p = reinterpret_cast<const bf16x8*>(d + v * 32);
        float s = hb[0];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bf16x8 u = p[q];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += b2f(u[j]) * h[8 * q + j];
        }
        y[v] = s;
    }
}

// Loss, with e = y - t and m = mask (1 when null):
//   L = [sum_v m_v e_v^2 + lambda sum_{x,y edges} (e_a - e_b)^2] / sum_v m_v
//       + sum_{z edges} w_v (e_v - e_{v+z})^2
// where the z-edge weight w_v is per node (edge v -- v + nx ny; null: lambda / sum m).
// gy = dL/dy.  Neighbour differences weight the errors that enlarge the residual of a
// start vector; weights proportional to the vertical conductances make the z term the
// energy of the error in the dominant couplings.
__global__ void k_loss(Vol o, const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ mask,
                       const float* __restrict__ wz, float inv_norm, float lambda, float* __restrict__ gy,
                       double* __restrict__ loss) {
    __shared__ float red[8];
    const long long nv = o.n();
    float ll = 0.f;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(v % o.nx), yy = (int)((v / o.nx) % o.ny), z = (int)(v / ((long long)o.nx * o.ny));
        const float m = mask ? mask[v] : 1.f;
        const float e = y[v] - t[v];
        float g = 2.f * m * e * inv_norm;
        float l = m * e * e * inv_norm;
        const long long sx = 1, sy = o.nx, sz = (long long)o.nx * o.ny;
        const float lx = lambda * inv_norm;
        if (lambda != 0.f) {
            if (x > 0) g += 2.f * lx * (e - (y[v - sx] - t[v - sx]));
            if (x < o.nx - 1) {
                const float d = e - (y[v + sx] - t[v + sx]);
                g += 2.f * lx * d;
                l += lx * d * d;
            }
            if (yy > 0) g += 2.f * lx * (e - (y[v - sy] - t[v - sy]));
            if (yy < o.ny - 1) {
                const float d = e - (y[v + sy] - t[v + sy]);
                g += 2.f * lx * d;
                l += lx * d * d;
            }
        }
        if (z > 0) {
            const float w = wz ? wz[v - sz] : lx;
            g += 2.f * w * (e - (y[v - sz] - t[v - sz]));
        }
        if (z < o.nz - 1) {
            const float w = wz ? wz[v] : lx;
            const float d = e - (y[v + sz] - t[v + sz]);
            g += 2.f * w * d;
            l += w * d * d;
        }
        gy[v] = g;
        ll += l;
    }
    for (int off = 32; off > 0; off >>= 1) ll += __shfl_xor(ll, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ll;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s2 = 0.f;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s2 += red[k];
        atomicAdd(loss, (double)s2);
    }
}

// Head backward: gpre[v][c] = gy_v h_c (d[v][c] > 0); gh[c] += sum_v gy_v d[v][c];
// ghb += sum_v gy_v.
__global__ void k_head_bwd(long long nv, const float* __restrict__ gy, const bf16* __restrict__ d,
                           const float* __restrict__ h, bf16* __restrict__ gpre, float* __restrict__ gh,
                           float* __restrict__ ghb) {
    __shared__ float red[33][8];
    float lh[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) lh[c] = 0.f;
    float lb = 0.f;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (long long)gridDim.x * blockDim.x) {
        const float g = gy[v];
        lb += g;
        const bf16x8* p = reinterpret_cast<const bf16x8*>(d + v * 32);
        bf16x8* q = reinterpret_cast<bf16x8*>(gpre + v * 32);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bf16x8 u = p[k];
            bf16x8 o;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float dv = b2f(u[j]);
                lh[8 * k + j] += g * dv;
                o[j] = f2b(dv > 0.f ? g * h[8 * k + j] : 0.f);
            }
            q[k] = o;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    auto wsum = [](float v) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const float s = wsum(lh[c]);
        if (lane == 0) red[c][wid] = s;
    }
    {
        const float s = wsum(lb);
        if (lane == 0) red[32][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < 33) {
        float s = 0.f;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[threadIdx.x][k];
        if (threadIdx.x < 32) atomicAdd(&gh[threadIdx.x], s);
        else atomicAdd(ghb, s);
    }
}

// Adam (Kingma & Ba 2015) on the fp32 master parameters; bias corrections c1 =
// 1 / (1 - b1^t), c2 = 1 / (1 - b2^t) from the host.
__global__ void k_adam(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i] * gscale;
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

// bf16 kernel copies of one conv layer's master weights w[27][cout][cin]:
// wf = w (forward), wt[t][ci][co] = w[26 - t][co][ci] (backward data).
__global__ void k_prep_weights(const float* __restrict__ w, int cout, int cin, int cin_pad, bf16* __restrict__ wf,
                               bf16* __restrict__ wt) {
    const long long n = 27LL * cout * cin;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(i % cin);
        const int co = (int)((i / cin) % cout);
        const int t = (int)(i / ((long long)cin * cout));
        const bf16 b = f2b(w[i]);
        wf[((size_t)t * cout + co) * cin_pad + ci] = b;
        if (wt) wt[((size_t)(26 - t) * cin + ci) * cout + co] = b;
    }
}

__global__ void k_f32_to_bf16(long long n, const float* __restrict__ a, bf16* __restrict__ b) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = f2b(a[i]);
}

__global__ void k_zero_f32(long long n, float* __restrict__ a) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a[i] = 0.f;
}

__global__ void k_zero_f64(long long n, double* __restrict__ a) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a[i] = 0.;
}

}  // namespace wn
