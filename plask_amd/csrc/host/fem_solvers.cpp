#include "fem_solvers.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>

namespace plask_amd {

static std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
static std::string fmt(const char* f, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

FemCore::FemCore(int dim, bool cylindrical, int device, std::string id_) : id(std::move(id_)), dim_(dim), cyl_(cylindrical) {
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    int rc = pf_ctx_create(&ctx_, device, 0);
    if (rc != PF_OK) throw std::runtime_error(id + ": no usable GPU device (pf_ctx_create " + std::to_string(rc) + ")");
}

FemCore::~FemCore() { pf_ctx_destroy(ctx_); }

void FemCore::check(int rc, const char* what) {
    if (rc == PF_OK) return;
    std::string msg = id + ": " + what + ": " + pf_last_error(ctx_);
    if (rc == PF_EBREAK || rc == PF_ENONFINITE || rc == PF_ENOCONV) throw ComputationError(msg);
    throw std::runtime_error(msg);
}

void FemCore::set_mesh(const std::vector<std::vector<double>>& axes) {
    if ((int)axes.size() != dim_) throw std::invalid_argument(id + ": mesh dimension does not match the solver");
    const double* a2 = dim_ == 3 ? axes[2].data() : nullptr;
    int64_t n2 = dim_ == 3 ? (int64_t)axes[2].size() : 1;
    check(pf_grid_set(ctx_, dim_, cyl_ ? PF_GEOM_CYLINDRICAL : PF_GEOM_CARTESIAN, axes[0].data(), axes[0].size(),
                      axes[1].data(), axes[1].size(), a2, n2),
          "mesh");
    N_ = pf_num_nodes(ctx_);
    E_ = pf_num_elements(ctx_);
    axes_ = axes;
    have_mesh_ = true;
    initialized_ = false;
}

void FemCore::set_dirichlet(const std::vector<int64_t>& nodes, const std::vector<double>& values) {
    if (nodes.size() != values.size()) throw std::invalid_argument("nodes and values differ in length");
    check(pf_dirichlet_set(ctx_, nodes.data(), values.data(), (int64_t)nodes.size()), "boundary conditions");
}

std::vector<double> FemCore::solution() {
    std::vector<double> x(N_);
    check(pf_get_solution(ctx_, x.data()), "solution");
    return x;
}

std::vector<double> FemCore::interp(const double* pts, int64_t n, double fill) {
    std::vector<double> out(n);
    check(pf_interp_nodes(ctx_, pts, n, fill, out.data()), "interpolation");
    return out;
}

std::vector<double> FemCore::postprocess(int what, const std::vector<double>& kh, const std::vector<double>& kv,
                                         bool fetch) {
    const bool given = !kh.empty();
    if (given && ((int64_t)kh.size() != E_ || (int64_t)kv.size() != E_))
        throw std::invalid_argument(id + ": conductivity arrays need one value per element");
    const size_t nc = what == PF_POST_JOULE ? 1 : (size_t)dim_;
    std::vector<double> out(fetch ? (size_t)E_ * nc : 0);
    check(pf_postprocess(ctx_, what, given ? kh.data() : nullptr, given ? kv.data() : nullptr,
                         fetch ? out.data() : nullptr),
          "post-processing");
    return out;
}

std::vector<double> FemCore::interp_elements(int what, const double* pts, int64_t n, double fill) {
    const size_t nc = what == PF_POST_JOULE ? 1 : (size_t)dim_;
    std::vector<double> out((size_t)n * nc);
    check(pf_interp_elements(ctx_, what, pts, n, fill, out.data()), "element interpolation");
    return out;
}

void FemCore::solve(double init_value, int norm) {
    pf_solve_opts o{};
    o.precision = iter.precision;
    o.precond = iter.precond ? iter.precond : 3;
    o.x0 = PF_X0_DEVICE;
    o.maxit = iter.maxit;
    o.rtol = iter.maxerr;
    o.inner_rtol = iter.inner_rtol;
    o.max_outer = 50;
    o.check_every = 32;
    o.norm = norm;
    o.init_value = init_value;
    pf_solve_stats st{};
    log(LOG_DETAIL, id + ": Iterating linear system");
    int rc = pf_solve(ctx_, &o, nullptr, &st);
    iter.iters = st.iters;
    iter.err = st.rel_resid;
    iter.seconds = st.seconds;
    log(LOG_DEBUG, fmt("%s: %d CG iterations in %d runs, %.3f ms, residual %g of %g", id.c_str(), st.iters, st.outer,
                       1e3 * st.seconds, st.rel_resid, st.bnorm));
    if (rc == PF_ENOCONV) {
        iter.converged = false;
        // iterative_matrix.hpp:299-314
        if (iter.noconv == 0)
            throw ComputationError(fmt("%s: failed to converge in %d iterations (error %g)", id.c_str(), iter.maxit,
                                       st.rel_resid));
        if (iter.noconv == 1)
            log(LOG_WARNING, fmt("%s: Failed to converge in %d iterations (error %g)", id.c_str(), iter.maxit, st.rel_resid));
        else
            log(LOG_DETAIL, fmt("%s: Did not converge yet in %d iterations (error %g)", id.c_str(), iter.maxit, st.rel_resid));
        return;
    }
    check(rc, "solve");
    iter.converged = true;
    log(LOG_DETAIL, fmt("%s: Converged after %d iterations (error %g)", id.c_str(), st.iters, st.rel_resid));
}

// ---------------------------------------------------------------------------
void ThermalCore::set_conductivity(const std::vector<double>& k0h, const std::vector<double>& k0v,
                                   const std::vector<double>& ah, const std::vector<double>& av) {
    if ((int64_t)k0h.size() != E_ || (int64_t)k0v.size() != E_ || (int64_t)ah.size() != E_ || (int64_t)av.size() != E_)
        throw std::invalid_argument(id + ": conductivity arrays need one value per element");
    k0h_ = k0h;
    k0v_ = k0v;
    ah_ = ah;
    av_ = av;
    tdep_ = std::any_of(ah.begin(), ah.end(), [](double a) { return a != 0.; }) ||
            std::any_of(av.begin(), av.end(), [](double a) { return a != 0.; });
    check(pf_coeffs_powerlaw(ctx_, PF_THERMAL, k0h.data(), k0v.data(), ah.data(), av.data(), E_), "conductivity");
    kcb_ = nullptr;
    if (!tdep_) check(pf_coeffs_set(ctx_, PF_THERMAL, k0h.data(), k0v.data(), E_), "conductivity");
}

void ThermalCore::set_conductivity_callback(std::function<void(const double*, double*, double*, int64_t)> cb) {
    kcb_ = std::move(cb);
}

void ThermalCore::set_heat(const std::vector<double>& q) {
    if (q.empty()) {
        check(pf_source_set(ctx_, nullptr, 0), "heat");
        return;
    }
    if ((int64_t)q.size() != E_) throw std::invalid_argument(id + ": heat array needs one value per element");
    check(pf_source_set(ctx_, q.data(), E_), "heat");
}

void ThermalCore::set_boundary_loads(const std::vector<int64_t>& flux_nodes, const std::vector<double>& flux,
                                     const std::vector<int64_t>& rad_nodes, const std::vector<double>& emissivity,
                                     const std::vector<double>& ambient) {
    if (!have_mesh_) throw std::runtime_error(id + ": no mesh");
    if (flux_nodes.size() != flux.size() || rad_nodes.size() != emissivity.size() || rad_nodes.size() != ambient.size())
        throw std::invalid_argument(id + ": boundary node and value lists differ in length");
    const double nan = std::nan("");
    if (flux_nodes.empty() && rad_nodes.empty()) {  // no conditions: nothing to evaluate per loop
        if (have_flux_ || have_rad_) check(pf_load_set(ctx_, nullptr, nullptr, 0), "boundary load");
        have_flux_ = have_rad_ = false;
        flux_.clear();
        emis_.clear();
        ramb_.clear();
        bc_elems_.clear();
        return;
    }
    have_flux_ = !flux_nodes.empty();
    have_rad_ = !rad_nodes.empty();
    flux_.assign(have_flux_ ? N_ : 0, nan);
    emis_.assign(have_rad_ ? N_ : 0, nan);
    ramb_.assign(have_rad_ ? N_ : 0, nan);
    std::vector<char> touched(N_, 0);
    auto node_ok = [&](int64_t n) {
        if (n < 0 || n >= N_) throw std::invalid_argument(id + ": boundary node out of range");
    };
    for (size_t q = 0; q < flux_nodes.size(); ++q) {
        const int64_t n = flux_nodes[q];
        node_ok(n);
        if (std::isnan(flux_[n])) flux_[n] = flux[q];
        touched[n] = 1;
    }
    for (size_t q = 0; q < rad_nodes.size(); ++q) {
        const int64_t n = rad_nodes[q];
        node_ok(n);
        if (std::isnan(emis_[n])) {
            emis_[n] = emissivity[q];
            ramb_[n] = ambient[q];
        }
        touched[n] = 1;
    }
    // elements touching a node under a condition (natural order, axis 0 fastest)
    bc_elems_.clear();
    const int64_t n0 = (int64_t)axes_[0].size(), n1 = (int64_t)axes_[1].size();
    const int64_t m0 = n0 - 1, m1 = n1 - 1;
    std::vector<char> ein(E_, 0);
    for (int64_t n = 0; n < N_; ++n) {
        if (!touched[n]) continue;
        const int64_t i = n % n0, j = (n / n0) % n1, k = dim_ == 3 ? n / (n0 * n1) : 0;
        const int64_t m2 = dim_ == 3 ? (int64_t)axes_[2].size() - 1 : 1;
        for (int64_t dk = (dim_ == 3 ? -1 : 0); dk <= 0; ++dk)
            for (int64_t dj = -1; dj <= 0; ++dj)
                for (int64_t di = -1; di <= 0; ++di) {
                    const int64_t ei = i + di, ej = j + dj, ek = k + dk;
                    if (ei < 0 || ei >= m0 || ej < 0 || ej >= m1 || ek < 0 || (dim_ == 3 && ek >= m2)) continue;
                    ein[ei + m0 * (ej + m1 * ek)] = 1;
                }
    }
    for (int64_t e = 0; e < E_; ++e)
        if (ein[e]) bc_elems_.push_back(e);
}

// The flux and radiation parts of setMatrix (therm2d.cpp:139-173, :232-240, :256-265;
// cylindrical :358-368, :394-406; therm3d.cpp:141-168, :242-249, :261-268): an element
// side whose nodes all carry a condition adds -len/2 * value per node (2D, cylindrical
// weights r -/+ len/6 along r and the edge radius on r = const sides) or -area/4 * value
// (3D; as in the reference the four values of any wall go to the element's local nodes
// 0-3).  Radiation's value is emissivity * SB * (T^4 - T_ambient^4) at the node's
// current temperature in 2D; in 3D the reference reads the temperature of the mesh node
// whose index is the wall node's local number (0-7), kept here.
void ThermalCore::boundary_load() {
    if ((have_flux_ && (int64_t)flux_.size() != N_) || (have_rad_ && (int64_t)emis_.size() != N_))
        throw std::runtime_error(id + ": boundary conditions were set for another mesh");
    const double SB = 5.670373e-8;  // plask/phys/constants.hpp:41
    std::vector<double> T;
    if (have_rad_) T = solution();
    std::vector<double> F(N_, 0.);
    std::vector<char> seen(N_, 0);
    std::vector<int64_t> nodes;
    auto add = [&](int64_t n, double v) {
        if (!seen[n]) {
            seen[n] = 1;
            nodes.push_back(n);
        }
        F[n] += v;
    };
    // condition value at node n; radiation uses the temperature of node nt
    auto value = [&](int kind, int64_t n, int64_t nt, double& v) -> bool {
        if (kind == 0) {
            if (!have_flux_ || std::isnan(flux_[n])) return false;
            v = flux_[n];
        } else {
            if (!have_rad_ || std::isnan(emis_[n])) return false;
            double a = ramb_[n], t = T[nt];
            a *= a;
            t *= t;
            v = emis_[n] * SB * (t * t - a * a);
        }
        return true;
    };
    const std::vector<double>& ax0 = axes_[0];
    const std::vector<double>& ax1 = axes_[1];
    const int64_t n0 = (int64_t)ax0.size(), n1 = (int64_t)ax1.size();
    const int64_t m0 = n0 - 1, m1 = n1 - 1;
    for (int kind = 0; kind < 2; ++kind) {
        if ((kind == 0 && !have_flux_) || (kind == 1 && !have_rad_)) continue;
        for (int64_t e : bc_elems_) {
            if (dim_ == 2) {
                const int64_t i = e % m0, j = e / m0;
                const int64_t lolo = i + n0 * j, uplo = lolo + 1, upup = uplo + n0, loup = lolo + n0;
                const double w = ax0[i + 1] - ax0[i], h = ax1[j + 1] - ax1[j];
                const double r = 0.5 * (ax0[i] + ax0[i + 1]);
                // sides in the reference's order: bottom, right, top, left
                const int64_t sa[4] = {lolo, uplo, upup, loup}, sb[4] = {uplo, upup, loup, lolo};
                const double len[4] = {w, h, w, h};
                for (int side = 0; side < 4; ++side) {
                    double va, vb;
                    if (!value(kind, sa[side], sa[side], va) || !value(kind, sb[side], sb[side], vb)) continue;
                    const double L = len[side];
                    auto weight = [&](int64_t a, int64_t b) {
                        if (!cyl_) return 1.;
                        if (side == 3) return ax0[i];      // left: r of the side
                        if (side == 1) return ax0[i + 1];  // right
                        return r + (a < b ? -L / 6. : L / 6.);
                    };
                    add(sa[side], -0.5e-6 * L * va * weight(sa[side], sb[side]));
                    add(sb[side], -0.5e-6 * L * vb * weight(sb[side], sa[side]));
                }
            } else {
                const int64_t i = e % m0, j = (e / m0) % m1, k = e / (m0 * m1);
                const std::vector<double>& ax2 = axes_[2];
                int64_t idx[8];
                for (int l = 0; l < 8; ++l)
                    idx[l] = (i + (l & 1)) + n0 * ((j + ((l >> 1) & 1)) + n1 * (k + (l >> 2)));
                const double dx = ax0[i + 1] - ax0[i], dy = ax1[j + 1] - ax1[j], dz = ax2[k + 1] - ax2[k];
                static const int walls[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 2, 4, 6},
                                                {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}};
                const double areas[3] = {dx * dy, dy * dz, dz * dx};
                for (int side = 0; side < 6; ++side) {
                    double v[4];
                    bool all = true;
                    // the reference passes the wall's local node number where the radiation
                    // term reads the temperature, i.e. the temperature of mesh node 0-7
                    for (int q = 0; q < 4 && all; ++q)
                        all = value(kind, idx[walls[side][q]], std::min<int64_t>(walls[side][q], N_ - 1), v[q]);
                    if (!all) continue;
                    for (int q = 0; q < 4; ++q) add(idx[q], -0.25e-12 * areas[side / 2] * v[q]);
                }
            }
        }
    }
    if (have_conv_)
        for (size_t q = 0; q < conv_nodes_.size(); ++q) add(conv_nodes_[q], conv_load_[q]);
    std::vector<double> vals(nodes.size());
    for (size_t q = 0; q < nodes.size(); ++q) vals[q] = F[nodes[q]];
    check(pf_load_set(ctx_, nodes.data(), vals.data(), (int64_t)nodes.size()), "boundary load");
}

// The convection part of setMatrix (setBoundaries with the convection functions):
// an element side whose two nodes (2D) or wall whose four nodes (3D) all carry a
// convection condition adds, as the reference does,
//   2D (therm2d.cpp:242-254): load 0.5e-6 len coeff_n ambient_n per node, matrix
//     (coeff_a + coeff_b) len / 6 on each diagonal and len / 12 off it (without the
//     1e-6 um -> m factor of the load: the reference's 2D convection matrix is 1e6 times
//     its load's scale, reproduced);
//   cylindrical (therm2d.cpp:370-392): load 0.125e-6 len (coeff_a + coeff_b)
//     (ambient_a + ambient_b) and the same matrix terms, each times r_edge on r = const
//     sides and (r_mid -/+ len/6) (diagonal and load) or r_mid (off-diagonal) along r;
//   3D (therm3d.cpp:250-259): load 0.25e-12 area coeff ambient per wall node and matrix
//     0.125e-12 area (coeff_i + coeff_j) / (9 | 18 | 36) for the same node, an edge or a
//     diagonal of the wall; as with the flux (therm3d.cpp:160-163) the wall's terms go to
//     the element's local nodes 0-3 (its lower face), reproduced.
void ThermalCore::set_convection(const std::vector<int64_t>& nodes, const std::vector<double>& coeff,
                                 const std::vector<double>& ambient) {
    if (!have_mesh_) throw std::runtime_error(id + ": no mesh");
    if (nodes.size() != coeff.size() || nodes.size() != ambient.size())
        throw std::invalid_argument(id + ": convection node and value lists differ in length");
    conv_nodes_.clear();
    conv_load_.clear();
    have_conv_ = !nodes.empty();
    if (!have_conv_) {
        if (conv_matrix_) check(pf_boundary_matrix_set(ctx_, nullptr, nullptr, nullptr, 0), "convection matrix");
        conv_matrix_ = false;
        return;
    }
    const double nan = std::nan("");
    std::vector<double> cc(N_, nan), ca(N_, nan);
    for (size_t q = 0; q < nodes.size(); ++q) {
        const int64_t n = nodes[q];
        if (n < 0 || n >= N_) throw std::invalid_argument(id + ": boundary node out of range");
        if (std::isnan(cc[n])) {
            cc[n] = coeff[q];
            ca[n] = ambient[q];
        }
    }
    auto has = [&](int64_t n) { return !std::isnan(cc[n]); };
    const std::vector<double>& ax0 = axes_[0];
    const std::vector<double>& ax1 = axes_[1];
    const int64_t n0 = (int64_t)ax0.size(), n1 = (int64_t)ax1.size();
    const int64_t m0 = n0 - 1, m1 = n1 - 1, m2 = dim_ == 3 ? (int64_t)axes_[2].size() - 1 : 1;
    std::vector<double> F(N_, 0.);
    std::vector<char> seen(N_, 0);
    std::vector<int64_t> rows, cols;
    std::vector<double> vals;
    auto load = [&](int64_t n, double v) {
        if (!seen[n]) {
            seen[n] = 1;
            conv_nodes_.push_back(n);
        }
        F[n] += v;
    };
    auto mat = [&](int64_t a, int64_t b, double v) {
        rows.push_back(std::max(a, b));
        cols.push_back(std::min(a, b));
        vals.push_back(v);
    };
    for (int64_t e = 0; e < E_; ++e) {
        if (dim_ == 2) {
            const int64_t i = e % m0, j = e / m0;
            const int64_t lolo = i + n0 * j, uplo = lolo + 1, upup = uplo + n0, loup = lolo + n0;
            if (!has(lolo) && !has(upup)) continue;  // every side has one of these two corners
            const double w = ax0[i + 1] - ax0[i], h = ax1[j + 1] - ax1[j];
            const double r = 0.5 * (ax0[i] + ax0[i + 1]);
            // sides in the reference's order: bottom, right, top, left
            const int64_t sa[4] = {lolo, uplo, upup, loup}, sb[4] = {uplo, upup, loup, lolo};
            const double len[4] = {w, h, w, h};
            for (int side = 0; side < 4; ++side) {
                const int64_t a = sa[side], b = sb[side];
                if (!has(a) || !has(b)) continue;
                const double L = len[side], csum = cc[a] + cc[b];
                if (!cyl_) {
                    load(a, 0.5e-6 * L * cc[a] * ca[a]);
                    load(b, 0.5e-6 * L * cc[b] * ca[b]);
                    mat(a, a, csum * L / 6.);
                    mat(b, b, csum * L / 6.);
                    mat(a, b, csum * L / 12.);
                } else {
                    auto wdiag = [&](int64_t p, int64_t o) {
                        if (side == 3) return ax0[i];      // left: r of the side
                        if (side == 1) return ax0[i + 1];  // right
                        return r + (p < o ? -L / 6. : L / 6.);
                    };
                    const double woff = side == 3 ? ax0[i] : side == 1 ? ax0[i + 1] : r;
                    const double fa = 0.125e-6 * L * csum * (ca[a] + ca[b]);
                    load(a, fa * wdiag(a, b));
                    load(b, fa * wdiag(b, a));
                    mat(a, a, csum * L / 6. * wdiag(a, b));
                    mat(b, b, csum * L / 6. * wdiag(b, a));
                    mat(a, b, csum * L / 12. * woff);
                }
            }
        } else {
            const int64_t i = e % m0, j = (e / m0) % m1, k = e / (m0 * m1);
            if (k >= m2) continue;
            int64_t idx[8];
            for (int l = 0; l < 8; ++l) idx[l] = (i + (l & 1)) + n0 * ((j + ((l >> 1) & 1)) + n1 * (k + (l >> 2)));
            if (!has(idx[0]) && !has(idx[7])) continue;  // every wall has one of these two corners
            const std::vector<double>& ax2 = axes_[2];
            const double dx = ax0[i + 1] - ax0[i], dy = ax1[j + 1] - ax1[j], dz = ax2[k + 1] - ax2[k];
            static const int walls[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 2, 4, 6},
                                            {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}};
            const double areas[3] = {dx * dy, dy * dz, dz * dx};
            for (int side = 0; side < 6; ++side) {
                const int* wl = walls[side];
                if (!has(idx[wl[0]]) || !has(idx[wl[1]]) || !has(idx[wl[2]]) || !has(idx[wl[3]])) continue;
                const double area = areas[side / 2];
                for (int p = 0; p < 4; ++p) {
                    const int64_t np_ = idx[wl[p]];
                    load(idx[p], 0.25e-12 * area * cc[np_] * ca[np_]);
                    for (int q = 0; q <= p; ++q) {
                        const int ij = p ^ q;
                        const double d = p == q ? 9. : (ij == 1 || ij == 2) ? 18. : 36.;
                        mat(idx[p], idx[q], 0.125e-12 * area * (cc[np_] + cc[idx[wl[q]]]) / d);
                    }
                }
            }
        }
    }
    conv_load_.resize(conv_nodes_.size());
    for (size_t q = 0; q < conv_nodes_.size(); ++q) conv_load_[q] = F[conv_nodes_[q]];
    check(pf_boundary_matrix_set(ctx_, rows.data(), cols.data(), vals.data(), (int64_t)vals.size()),
          "convection matrix");
    conv_matrix_ = !vals.empty();
}

// onInitialize: temperatures = inittemp, loopno = 0 (therm3d.cpp:71-80)
void ThermalCore::initialize() {
    if (!have_mesh_) throw std::runtime_error(id + ": no mesh");
    std::vector<double> t(N_, inittemp);
    check(pf_solution_set(ctx_, t.data()), "initial temperature");
    loopno = 0;
    initialized_ = true;
}

void ThermalCore::update_conductivity() {
    if (kcb_) {
        std::vector<double> tavg(E_), kh(E_), kv(E_);
        check(pf_element_means(ctx_, tavg.data()), "element temperatures");
        kcb_(tavg.data(), kh.data(), kv.data(), E_);
        check(pf_coeffs_set(ctx_, PF_THERMAL, kh.data(), kv.data(), E_), "conductivity");
    } else if (tdep_) {
        check(pf_coeffs_update(ctx_), "conductivity");
    }
}

double ThermalCore::compute(int loops) {
    if (!initialized_) initialize();
    log(LOG_INFO, id + ": Running thermal calculations");
    if (heat_src_) check(pf_source_from(ctx_, heat_src_->context(), PF_POST_JOULE), "heat from the linked solver");
    int loop = 0;
    double err = 0.;
    toterr = 0.;
    check(pf_solution_change(ctx_, nullptr, nullptr), "loop start");  // T_prev = T
    do {
        // setMatrix: element conductivities at the current element temperatures
        update_conductivity();
        // boundary loads at the loop's start temperatures (radiation changes every loop)
        if (have_rad_ || ((have_flux_ || have_conv_) && loop == 0)) boundary_load();
        solve(inittemp);
        check(pf_solution_change(ctx_, &err, &maxT), "loop error");
        if (err > toterr) toterr = err;
        ++loopno;
        ++loop;
        log(LOG_RESULT, fmt("%s: Loop %d(%d): max(T) = %.3f K, error = %g K", id.c_str(), loop, loopno, maxT, err));
    } while ((!iter.converged || err > maxerr) && (loops == 0 || loop < loops));
    return toterr;
}

// ---------------------------------------------------------------------------
// onInitialize: temperatures = inittemp, elapstime = 0 (femT3d.cpp:71-80)
void DynamicCore::initialize() {
    ThermalCore::initialize();
    elapstime = 0.;
}

// setMatrix (femT3d.cpp:140-279): conductivities and capacities at the element
// temperatures, A = theta K + C on the device
void DynamicCore::set_matrix() {
    log(LOG_DETAIL, id + ": Setting up matrix system");
    update_conductivity();
    if (!ccb_) throw std::runtime_error(id + ": no heat capacity");
    std::vector<double> tavg(E_), c(E_);
    check(pf_element_means(ctx_, tavg.data()), "element temperatures");
    ccb_(tavg.data(), c.data(), E_);
    check(pf_capacity_set(ctx_, c.data(), E_, methodparam), "heat capacity");
}

// compute(time) (femT3d.cpp:282-331): steps of timestep from t = 0 while t < time + timestep/2,
// rebuilding the matrices every rebuildfreq steps and logging max(T) every logfreq steps
double DynamicCore::compute(double time) {
    if (!initialized_) initialize();
    if (!(timestep > 0.)) throw std::invalid_argument(id + ": timestep must be positive");
    log(LOG_INFO, id + ": Running thermal calculations");
    set_matrix();
    pf_solve_opts o{};
    o.precision = iter.precision;
    o.precond = iter.precond ? iter.precond : 3;
    o.maxit = iter.maxit;
    o.rtol = iter.maxerr;
    o.inner_rtol = iter.inner_rtol;
    o.max_outer = 50;
    o.check_every = 32;
    // each step solved as the reference's back-substitution would: the step's start residual
    // reduced by maxerr, floored at 1e-4 ||B|| (X_init = 0) once the steps become small
    o.norm = PF_NORM_START;
    o.init_value = 0.;
    unsigned long r = (unsigned long)rebuildfreq, l = (unsigned long)logfreq;
    time += timestep / 2.;
    for (double t = 0.; t < time; t += timestep) {
        if (rebuildfreq && r == 0) {
            set_matrix();
            r = (unsigned long)rebuildfreq;
        }
        pf_solve_stats st{};
        int rc = pf_dynamic_step(ctx_, &o, held, &st);
        iter.iters = st.iters;
        iter.err = st.rel_resid;
        iter.seconds = st.seconds;
        if (rc == PF_ENOCONV) {
            iter.converged = false;
            if (iter.noconv == 0)
                throw ComputationError(fmt("%s: failed to converge in %d iterations (error %g)", id.c_str(), iter.maxit,
                                           st.rel_resid));
            if (iter.noconv == 1)
                log(LOG_WARNING, fmt("%s: Failed to converge in %d iterations (error %g)", id.c_str(), iter.maxit,
                                     st.rel_resid));
        } else {
            check(rc, "time step");
            iter.converged = true;
        }
        if (logfreq && l == 0) {
            check(pf_solution_change(ctx_, nullptr, &maxT), "max temperature");
            log(LOG_RESULT, fmt("%s: Time %.2f ns: max(T) = %.3f K", id.c_str(), elapstime, maxT));
            l = (unsigned long)logfreq;
        }
        --r;
        --l;
        elapstime += timestep;
        ++loopno;
    }
    elapstime -= timestep;
    return 0.;
}

// ---------------------------------------------------------------------------
void ElectricalCore::set_conductivity(const std::vector<double>& sh, const std::vector<double>& sv) {
    if ((int64_t)sh.size() != E_ || (int64_t)sv.size() != E_)
        throw std::invalid_argument(id + ": conductivity arrays need one value per element");
    sh_ = sh;
    sv_ = sv;
}

void ElectricalCore::set_junctions(const std::vector<int64_t>& elems, int ncorner, const std::vector<int64_t>& bottom,
                                   const std::vector<int64_t>& top, const std::vector<int>& junction,
                                   const std::vector<double>& heights) {
    check(pf_junctions_set(ctx_, (int64_t)elems.size(), ncorner, elems.data(), bottom.data(), top.data(),
                           junction.data(), (int)heights.size(), heights.data()),
          "junctions");
    jelems_ = elems;
    jcond_.assign(elems.size(), start_cond_);
    jcond_h_.assign(elems.size(), start_cond_h_);
}

void ElectricalCore::set_junction_params(const std::vector<double>& beta, const std::vector<double>& js) {
    check(pf_junction_params(ctx_, beta.data(), js.data(), (int64_t)beta.size()), "junction parameters");
}

void ElectricalCore::set_flags(const std::vector<unsigned char>& select, const std::vector<unsigned char>& noheat) {
    check(pf_element_flags_set(ctx_, select.empty() ? nullptr : select.data(), noheat.empty() ? nullptr : noheat.data()),
          "element flags");
    have_flags_ = true;
}

// onInitialize: potentials = 0, loopno = 0, junction conductivity = start value (electr2d.cpp:171-187)
void ElectricalCore::initialize() {
    if (!have_mesh_) throw std::runtime_error(id + ": no mesh");
    std::vector<double> v(N_, 0.);
    check(pf_solution_set(ctx_, v.data()), "initial potential");
    loopno = 0;
    jcond_.assign(jelems_.size(), start_cond_);
    jcond_h_.assign(jelems_.size(), start_cond_h_);
    coeffs_live_ = jcond_on_device_ = false;
    initialized_ = true;
}

void ElectricalCore::link_temperature(FemCore* src, const std::vector<double>& s0h, const std::vector<double>& s0v,
                                      const std::vector<double>& ah, const std::vector<double>& av) {
    if (!src) {
        temp_src_ = nullptr;
        return;
    }
    if ((int64_t)s0h.size() != E_ || (int64_t)s0v.size() != E_ || (int64_t)ah.size() != E_ ||
        (int64_t)av.size() != E_)
        throw std::invalid_argument(id + ": conductivity law arrays need one value per element");
    check(pf_coeffs_powerlaw(ctx_, PF_ELECTRICAL, s0h.data(), s0v.data(), ah.data(), av.data(), E_),
          "conductivity law");
    temp_src_ = src;
}

// the junction conductivities the device holds (saveConductivities, electr2d.cpp:355-361)
void ElectricalCore::fetch_jcond() {
    if (!jelems_.empty()) {
        std::vector<double> ch(E_), cv(E_);
        check(pf_get_coeffs(ctx_, ch.data(), cv.data()), "conductivity");
        for (size_t q = 0; q < jelems_.size(); ++q) {
            jcond_[q] = cv[jelems_[q]];
            jcond_h_[q] = ch[jelems_[q]];
        }
    }
    jcond_on_device_ = false;
}

double ElectricalCore::compute(int loops) {
    if (!initialized_) initialize();
    if ((int64_t)sh_.size() != E_) throw std::runtime_error(id + ": conductivities not set");
    log(LOG_INFO, id + ": Running electrical calculations");
    // loadConductivities (electr2d.cpp:331-353): materials, then the saved junction values
    if (temp_src_ && coeffs_live_) {
        // on the device: the law at the linked temperatures; junction elements keep theirs
        check(pf_coeffs_update_from(ctx_, temp_src_->context()), "conductivity at the linked temperatures");
    } else {
        if (jcond_on_device_) fetch_jcond();
        std::vector<double> kh = sh_, kv = sv_;
        for (size_t q = 0; q < jelems_.size(); ++q) {
            kh[jelems_[q]] = jcond_h_[q];
            kv[jelems_[q]] = (std::isnan(jcond_[q]) || std::abs(jcond_[q]) < 1e-16) ? 1e-16 : jcond_[q];
        }
        check(pf_coeffs_set(ctx_, PF_ELECTRICAL, kh.data(), kv.data(), E_), "conductivity");
        coeffs_live_ = true;
    }
    const bool noactive = jelems_.empty();
    const double minj = 100e-7;  // electr2d.cpp:391
    unsigned loop = 0;
    double err = 0.;
    toterr = 0.;
    do {
        if (loopno != 0) {
            if (law_) {
                const int64_t nq = (int64_t)jelems_.size();
                std::vector<double> U(nq), jy(nq), ch(nq, 0.), cv(nq, 0.);
                check(pf_junction_state(ctx_, U.data(), jy.data(), nq), "junction state");
                law_(U.data(), jy.data(), ch.data(), cv.data(), nq);
                check(pf_junction_cond_set(ctx_, ch.data(), cv.data(), nq, stable ? 1 : 0), "junction conductivity");
            } else {
                check(pf_junction_update(ctx_, stable ? 1 : 0), "junction update");
            }
        }
        // every loop solves the system of its junction conductivities to the tolerance, from
        // the previous loop's potential (the reference solves each loop directly in 2D)
        solve(0., PF_NORM_START);
        double dj2 = 0., j2 = 0.;
        check(pf_currents(ctx_, &dj2, &j2), "currents");
        double mcur = std::sqrt(j2);
        err = 100. * std::sqrt(dj2) / std::max(mcur, minj);
        if ((loop != 0 || mcur >= minj) && err > toterr) toterr = err;
        maxcur = mcur;
        ++loopno;
        ++loop;
        log(LOG_RESULT, fmt("%s: Loop %d(%d): max(j%s) = %g kA/cm2, error = %g%%", id.c_str(), loop, loopno,
                            noactive ? "" : "@junc", mcur, err));
    } while ((!iter.converged || err > maxerr) && (loops == 0 || loop < (unsigned)loops));
    // saveConductivities (electr2d.cpp:355-361): linked, the junction values stay on the device
    jcond_on_device_ = true;
    if (!temp_src_) fetch_jcond();
    return toterr;
}

std::vector<double> ElectricalCore::currents() {
    std::vector<double> c(E_ * dim_);
    check(pf_get_currents(ctx_, c.data(), nullptr), "currents");
    return c;
}

std::vector<double> ElectricalCore::heats() {
    std::vector<double> h(E_);
    check(pf_get_currents(ctx_, nullptr, h.data()), "heats");
    return h;
}

std::vector<double> ElectricalCore::conductivity_h() {
    std::vector<double> k(E_);
    check(pf_get_coeffs(ctx_, k.data(), nullptr), "conductivity");
    return k;
}

std::vector<double> ElectricalCore::conductivity_v() {
    std::vector<double> k(E_);
    check(pf_get_coeffs(ctx_, nullptr, k.data()), "conductivity");
    return k;
}

}  // namespace plask_amd
