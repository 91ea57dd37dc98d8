// vf_host.hip -- host orchestration of the MI355X solve-and-sensitivity path and the C ABI
// declared in include/voxelfem.h.
//
// The simulator (vf_sim) keeps the fine grid state resident in HBM: densities, Young's moduli in
// FP64 and FP32, the element matrix K0, Dirichlet masks/lists and the load data. The multigrid
// solver (vf_mg) owns the coarsened hierarchy (Galerkin element matrices, 27-point block rows,
// the FP64 inverse of the reduced coarsest operator) and all level work buffers, and runs the
// V-cycle / full multigrid preconditioner and the PCG iteration entirely on the device; the host
// reads back one scalar per PCG iteration for the stopping test.
#include "vf_internal.hpp"
#include "vf_planar.hpp"
#include "../../include/voxelfem.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>

using namespace vf;

namespace {

thread_local std::string g_err;
constexpr int64_t LAYER_NONE = INT64_MAX;
// Fine-level kernel selection: 0 = chosen per solver and precision by timing both families on the
// solver's own grid (autotuned), 1 = register-gather kernels, 2 = LDS-tiled structured kernels (where
// the element matrix has the structure they need).
int g_kernel_variant = 0;

void sync(hipStream_t s) { VF_HIP(hipStreamSynchronize(s)); }

template <typename T>
void h2d(T *dst, const T *src, size_t n, hipStream_t s) {
    // The sources are ordinary (pageable) host buffers, often temporaries; the copy is complete when this
    // returns, so they may be reused or freed right away.
    if (n) {
        VF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
        VF_HIP(hipStreamSynchronize(s));
    }
}
template <typename T>
void d2h(T *dst, const T *src, size_t n, hipStream_t s) {
    if (n) VF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <typename T>
void d2d(T *dst, const T *src, size_t n, hipStream_t s) {
    if (n && dst != src) VF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

// ---- element matrices (host, once per material/stretch change) -------------------------------
// K0 = int eps(phi_a e_c) : C : eps(phi_b e_d) over the box element (TensorProductSimulator.hh:67-80)
// evaluated in closed form from the 1D integrals of the Q1 basis on [0, h]:
//   mass M = h [[1/3,1/6],[1/6,1/3]], stiffness D = (1/h) [[1,-1],[-1,1]], mixed G = [[-1/2,-1/2],[1/2,1/2]].
void element_K0(double young, double nu, const double h[3], double K0[576]) {
    const double lam = young * nu / ((1 + nu) * (1 - 2 * nu));
    const double mu = young / (2 * (1 + nu));
    auto Mm = [&](int d, int a, int b) { return h[d] * (a == b ? 1.0 / 3.0 : 1.0 / 6.0); };
    auto Dm = [&](int d, int a, int b) { return (a == b ? 1.0 : -1.0) / h[d]; };
    auto Gm = [&](int a, int b) { (void)b; return a ? 0.5 : -0.5; };  // int N_a' N_b
    // I(p,q)[a][b] = int d_p phi_a d_q phi_b
    auto I = [&](int p, int q, const int ia[3], const int ib[3]) {
        double v = 1;
        for (int d = 0; d < 3; ++d) {
            if (d == p && d == q) v *= Dm(d, ia[d], ib[d]);
            else if (d == p) v *= Gm(ia[d], ib[d]);
            else if (d == q) v *= Gm(ib[d], ia[d]);
            else v *= Mm(d, ia[d], ib[d]);
        }
        return v;
    };
    for (int a = 0; a < 8; ++a) {
        const int ia[3] = {(a >> 2) & 1, (a >> 1) & 1, a & 1};
        for (int b = 0; b < 8; ++b) {
            const int ib[3] = {(b >> 2) & 1, (b >> 1) & 1, b & 1};
            for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d) {
                    double v = lam * I(c, d, ia, ib) + mu * I(d, c, ia, ib);
                    if (c == d) v += mu * (I(0, 0, ia, ib) + I(1, 1, ia, ib) + I(2, 2, ia, ib));
                    K0[(3 * a + c) * 24 + 3 * b + d] = v;
                }
        }
    }
    for (int i = 0; i < 24; ++i)  // exact symmetry
        for (int j = 0; j < i; ++j) K0[j * 24 + i] = K0[i * 24 + j];
}

// phi[fi][lf][lc]: coarse Q1 shape function lc evaluated at fine node lf of child fi
void child_phi(double phi[512]) {
    for (int fi = 0; fi < 8; ++fi)
        for (int lf = 0; lf < 8; ++lf)
            for (int lc = 0; lc < 8; ++lc) {
                double v = 1;
                for (int d = 0; d < 3; ++d) {
                    const int sh = 2 - d;
                    const double x = 0.5 * (((fi >> sh) & 1) + ((lf >> sh) & 1));
                    v *= ((lc >> sh) & 1) ? x : 1 - x;
                }
                phi[fi * 64 + lf * 8 + lc] = v;
            }
}

// out += P_fi^T K P_fi with P[3 lf + c][3 lc + c] = phi[fi][lf][lc]
void coarsen_into(const double *phi, int fi, const double *K, double *out) {
    double T[576];
    for (int a = 0; a < 24; ++a)
        for (int j = 0; j < 8; ++j)
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int i = 0; i < 8; ++i) s += K[a * 24 + 3 * i + d] * phi[fi * 64 + i * 8 + j];
                T[a * 24 + 3 * j + d] = s;
            }
    for (int i2 = 0; i2 < 8; ++i2)
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < 24; ++b) {
                double s = 0;
                for (int i = 0; i < 8; ++i) s += phi[fi * 64 + i * 8 + i2] * T[(3 * i + c) * 24 + b];
                out[(3 * i2 + c) * 24 + b] += s;
            }
}

}  // namespace

// =================================================================================================
struct vf_sim {
    hipStream_t stream = nullptr;
    // planar (2D) simulator: when set, every entry point works on it and the 3D state below stays empty
    std::unique_ptr<vf::planar::PSim> p2;
    int64_t ne[3];
    double lo[3], hi[3], h[3];
    Grid g;
    double young = 1, poisson = 0;  // ETensor default E=1, nu=0 (TensorProductSimulator.hh:2114)
    int law = VF_SIMP;
    double E0 = 1, Emin = 1e-4, gamma = 3, q = 3;  // TensorProductSimulator.hh:2137-2142
    double grav[3] = {0, 0, 0};
    double K0h[576];
    DBuf<double> K0;
    DBuf<float> K032;
    DBuf<double> Ftab;  // structured element matrix (see update_K0)
    DBuf<float> Ftab32;
    bool structured = false;
    DBuf<double> Wtab;  // Walsh-basis element matrix entries (see update_K0)
    DBuf<float> Wtab32;
    bool walsh = false;
    DBuf<double> rho, E;
    DBuf<float> E32;
    std::vector<uint8_t> hmask;
    std::vector<double> hvals;
    DBuf<uint8_t> dmask;
    DBuf<int> dnodes;
    DBuf<uint8_t> dnmask;
    DBuf<double> dvals;
    int ndir = 0;
    std::vector<double> hforce;
    DBuf<double> force;
    bool hasForce = false;
    int64_t firstMasked = LAYER_NONE, firstDetached = LAYER_NONE;
    double maskHeight = std::numeric_limits<double>::infinity();
    DBuf<double> s_in, s_out, s_aux, partials, scal;
    DBuf<float> s32a, s32b;
    uint64_t bcVersion = 0;    // bumped whenever the constraint masks change
    uint64_t version = 0;      // bumped whenever E or K0 changes
    uint64_t baseVersion = 0;  // the same, except when layers are removed from the top of the mask

    size_t nn() const { return (size_t)g.nnodes; }
    size_t nel() const { return (size_t)g.nelems; }

    void refresh_grid() {
        g.ydet = (int)std::min<int64_t>(firstDetached, g.nn[1]);
        g.ymask = (int)std::min<int64_t>(firstMasked, g.ne[1]);
    }
    void update_K0() {
        element_K0(young, poisson, h, K0h);
        float f[576];
        for (int i = 0; i < 576; ++i) f[i] = (float)K0h[i];
        h2d(K0.data(), K0h, 576, stream);
        h2d(K032.data(), f, 576, stream);
        // Structured form used by the LDS-tiled fine kernels: K0[(l,c),(l^delta,d)] equals
        // F[delta][c][d] = K0[(0,c),(delta,d)] times s_c(l) s_d(l) for c != d (s_a(l) = -1 or +1
        // for the low or high side of axis a), which holds for box elements of an orthotropic tensor.
        double F[128] = {0};
        float F32[128] = {0};
        for (int dl = 0; dl < 8; ++dl)
            for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d) F[16 * dl + 3 * c + d] = K0h[c * 24 + 3 * dl + d];
        double kmax = 0, dev = 0;
        for (int i = 0; i < 576; ++i) kmax = std::max(kmax, std::fabs(K0h[i]));
        for (int l = 0; l < 8; ++l)
            for (int dl = 0; dl < 8; ++dl)
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d) {
                        const int sc = ((l >> (2 - c)) & 1) ? 1 : -1, sd = ((l >> (2 - d)) & 1) ? 1 : -1;
                        const double sg = c == d ? 1.0 : double(sc * sd);
                        const double v = K0h[(3 * l + c) * 24 + 3 * (l ^ dl) + d];
                        dev = std::max(dev, std::fabs(v - sg * F[16 * dl + 3 * c + d]));
                    }
        structured = dev <= 1e-13 * kmax;
        for (int i = 0; i < 128; ++i) F32[i] = (float)F[i];
        h2d(Ftab.data(), F, 128, stream);
        h2d(Ftab32.data(), F32, 128, stream);
        // Walsh basis: Khat = W K0 W^T with W[k][m] = prod_a s_a(m)^{k_a}; the element-centric kernel
        // uses Khat[(k,d),(k,d)] and Khat[(k,d),(k^e_d^e_c,c)] (c != d, k_d != k_c), divided by 64.
        double Kh[576] = {0};
        auto Wkm = [](int k, int m) {
            int sgn = 1;
            for (int a = 0; a < 3; ++a)
                if ((k >> (2 - a)) & 1) sgn *= ((m >> (2 - a)) & 1) ? 1 : -1;
            return sgn;
        };
        for (int k = 0; k < 8; ++k)
            for (int l = 0; l < 8; ++l)
                for (int d = 0; d < 3; ++d)
                    for (int c = 0; c < 3; ++c) {
                        double acc = 0;
                        for (int m = 0; m < 8; ++m)
                            for (int q = 0; q < 8; ++q)
                                acc += Wkm(k, m) * K0h[(3 * m + d) * 24 + 3 * q + c] * Wkm(l, q);
                        Kh[(3 * k + d) * 24 + 3 * l + c] = acc;
                    }
        double C[128] = {0}, khmax = 0, off = 0;
        for (int i = 0; i < 576; ++i) khmax = std::max(khmax, std::fabs(Kh[i]));
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < 3; ++d)
                for (int l = 0; l < 8; ++l)
                    for (int c = 0; c < 3; ++c) {
                        const int bd = 1 << (2 - d), bc = 1 << (2 - c);
                        const bool kd = (k & bd) != 0, kc = (k & bc) != 0;
                        const bool used = (c == d && l == k) || (c != d && kd != kc && l == (k ^ bd ^ bc));
                        const double v = Kh[(3 * k + d) * 24 + 3 * l + c];
                        if (used) C[16 * k + 3 * d + c] = v / 64.0;
                        else off = std::max(off, std::fabs(v));
                    }
        walsh = off <= 1e-13 * khmax;
        float C32[128];
        for (int i = 0; i < 128; ++i) C32[i] = (float)C[i];
        h2d(Wtab.data(), C, 128, stream);
        h2d(Wtab32.data(), C32, 128, stream);
        sync(stream);
    }
    void update_E() {
        refresh_grid();
        launch_young(g, rho.data(), E.data(), E32.data(), law, E0, Emin, gamma, q, maskHeight < hi[1] ? 1 : 0, stream);
        ++version;
        ++baseVersion;
    }
    // setFabricationMaskHeight (TensorProductSimulator.hh:286-310)
    void set_mask_height(double H, bool updateYoung) {
        if (H < 0 || H > hi[1])
            throw std::runtime_error("Fabrication height (" + std::to_string(H) + ") has to be in between 0 and " +
                                     std::to_string(hi[1]));
        maskHeight = H;
        firstMasked = (int64_t)std::ceil(maskHeight / h[1] - 1e-10);
        firstDetached = firstMasked + 1;
        refresh_grid();
        if (updateYoung) update_E();
    }
    // decrementFabricationMaskHeightByLayer (TensorProductSimulator.hh:311-324): mask `count` more element layers
    // from the top, zeroing their moduli in place (the moduli update_E would give at the lower height). Only
    // `version` changes, so a hierarchy built before can be updated by bands.
    void remove_top_layers(int64_t count) {
        maskHeight -= (double)count * h[1];
        firstMasked -= count;
        firstDetached = firstMasked + 1;
        refresh_grid();
        launch_zero_young_layers(g, (int)firstMasked, (int)(firstMasked + count), E.data(), E32.data(), stream);
        ++version;
    }
    double *stage_in(size_t n) {
        s_in.resize(std::max(s_in.size(), n));
        return s_in.data();
    }
    double *stage_out(size_t n) {
        s_out.resize(std::max(s_out.size(), n));
        return s_out.data();
    }
    double *stage_aux(size_t n) {
        s_aux.resize(std::max(s_aux.size(), n));
        return s_aux.data();
    }
    double *red_partials(size_t n) {
        partials.resize(std::max(partials.size(), std::max<size_t>(n, 64)));
        return partials.data();
    }
    double *scalars() {
        if (scal.empty()) scal.resize(16);
        return scal.data();
    }
    // deep copy of the simulator's state (device arrays copied on this simulator's stream) bound to `st`
    std::unique_ptr<vf_sim> clone(hipStream_t st) const {
        auto c = std::make_unique<vf_sim>();
        c->stream = st;
        std::copy(ne, ne + 3, c->ne);
        std::copy(lo, lo + 3, c->lo);
        std::copy(hi, hi + 3, c->hi);
        std::copy(h, h + 3, c->h);
        c->g = g;
        c->young = young; c->poisson = poisson; c->law = law;
        c->E0 = E0; c->Emin = Emin; c->gamma = gamma; c->q = q;
        std::copy(grav, grav + 3, c->grav);
        std::copy(K0h, K0h + 576, c->K0h);
        c->structured = structured;
        c->walsh = walsh;
        c->hmask = hmask;
        c->hvals = hvals;
        c->ndir = ndir;
        c->hforce = hforce;
        c->hasForce = hasForce;
        c->firstMasked = firstMasked;
        c->firstDetached = firstDetached;
        c->maskHeight = maskHeight;
        c->version = version;
        c->baseVersion = baseVersion;
        auto dup = [&](auto &dst, const auto &src) {
            dst.resize(src.size());
            d2d(dst.data(), src.data(), src.size(), stream);
        };
        dup(c->K0, K0); dup(c->K032, K032); dup(c->Ftab, Ftab); dup(c->Ftab32, Ftab32); dup(c->Wtab, Wtab);
        dup(c->Wtab32, Wtab32); dup(c->rho, rho); dup(c->E, E); dup(c->E32, E32); dup(c->dmask, dmask);
        dup(c->dnodes, dnodes); dup(c->dnmask, dnmask); dup(c->dvals, dvals); dup(c->force, force);
        sync(stream);
        return c;
    }
    double read_scalar(const double *dev) {
        double v;
        d2h(&v, dev, 1, stream);
        sync(stream);
        return v;
    }
};

namespace {

struct LevelData {
    Grid g;
    double h[3];
    std::vector<uint8_t> hmask;
    DBuf<uint8_t> dmask;
    DBuf<int> dnodes;
    DBuf<uint8_t> dnmask;
    int ndir = 0;
    DBuf<double> x, b, r;
    DBuf<float> x32, b32, r32;
    DBuf<double> blocks;
    DBuf<float> blocks32;
    DBuf<double> Ke;
};

void upload_dirichlet(const std::vector<uint8_t> &mask, DBuf<uint8_t> &dmask, DBuf<int> &dnodes,
                      DBuf<uint8_t> &dnmask, int &ndir, hipStream_t s) {
    std::vector<int> nodes;
    std::vector<uint8_t> m;
    for (size_t n = 0; n < mask.size(); ++n)
        if (mask[n] & 7) { nodes.push_back((int)n); m.push_back(mask[n] & 7); }
    ndir = (int)nodes.size();
    dmask.resize(mask.size());
    h2d(dmask.data(), mask.data(), mask.size(), s);
    dnodes.resize(std::max<size_t>(nodes.size(), 1));
    dnmask.resize(std::max<size_t>(m.size(), 1));
    h2d(dnodes.data(), nodes.data(), nodes.size(), s);
    h2d(dnmask.data(), m.data(), m.size(), s);
    sync(s);
}

}  // namespace

template <typename T>
void sim_apply_K(vf_sim *sim, const T *u, T *y, hipStream_t s) {
    const T *E, *K0, *F;
    const T *Wc;
    if constexpr (std::is_same<T, double>::value) {
        E = sim->E.data(); K0 = sim->K0.data(); F = sim->Ftab.data(); Wc = sim->Wtab.data();
    } else {
        E = sim->E32.data(); K0 = sim->K032.data(); F = sim->Ftab32.data(); Wc = sim->Wtab32.data();
    }
    if (sim->walsh && (g_kernel_variant == 0 || g_kernel_variant >= 3))
        launch_fine_apply_walsh<T>(sim->g, Wc, E, u, y, nullptr, nullptr, AM_SET, s);
    else if (sim->structured && g_kernel_variant != 1)
        launch_fine_apply_tiled<T>(sim->g, F, E, u, y, nullptr, nullptr, AM_SET, s);
    else
        launch_fine_apply<T>(sim->g, K0, E, u, y, nullptr, nullptr, AM_SET, s);
}

namespace {
// owner of a captured graph and its executable instance
struct GraphGuard {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphGuard() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};
}  // namespace

struct StackedBatch;  // a stacked layer-by-layer batch (see lbl_stacked_batch)

struct vf_mg {
    vf_sim *sim;
    int L;
    std::unique_ptr<vf::planar::PMG> p2;  // planar hierarchy (the 3D state below stays empty)
    std::vector<std::unique_ptr<LevelData>> lv;  // lv[0] mirrors the fine simulator
    double phi[512];
    double Chat[8 * 576];
    DBuf<double> dChat, dphi, dM2, dKe0, dCw;
    DBuf<float> dChat32, dCw32;
    bool symmetricGS = true;
    bool useGraphs = false;  // replay the steady-state PCG iteration from a captured HIP graph
    // the last solve's iteration graph, destroyed during the next solve's capture (while the device is busy)
    // rather than with the device idle at the end of its own solve
    GraphGuard retiredGraph;
    void retire_graph() {
        if (retiredGraph.exec) VF_HIP(hipGraphExecDestroy(retiredGraph.exec));
        if (retiredGraph.graph) VF_HIP(hipGraphDestroy(retiredGraph.graph));
        retiredGraph.exec = nullptr;
        retiredGraph.graph = nullptr;
    }
    bool l1Blocks = false;  // level 1 held as assembled block rows (set by update_stiffness)
    // block rows of the current stiffness in FP64 / FP32, assembled on first use by a cycle of that precision
    bool blocksReady[2] = {false, false};
    // coarsest operator
    DBuf<double> Ainv, AinvT, invScratch, ctmp_in, ctmp_out, band;  // X = L^-1 and X^T (see prepare_coarsest, launch_factor)
    DBuf<int> cholFlag;
    // the coarsest factorization runs on its own stream, overlapping the block assembly and the first
    // cycle's descent; the main stream waits for it at the first coarse solve (finish_factor)
    hipStream_t fstream = nullptr;
    hipEvent_t evGalerkin = nullptr, evFactor = nullptr;
    bool factorPending = false;
    // The Galerkin chain of levels >= 2 and the block rows assembled from it run on a third stream (default
    // priority), after the moduli (evModuli): the first cycle's fine-level and level-1 work, which needs only
    // the moduli, overlaps it. The solver's stream waits for evHier before its first use of level >= 2 data
    // (wait_hier); hierPending marks that wait as due.
    hipStream_t gstream = nullptr;
    hipEvent_t evModuli = nullptr, evHier = nullptr;
    bool hierPending = false;
    bool overlapSetup = false;  // queue the Galerkin chain on gstream (measured no faster on the full-size cantilever)
    void ensure_gstream() {
        if (gstream) return;
        VF_HIP(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
        VF_HIP(hipEventCreateWithFlags(&evModuli, hipEventDisableTiming));
        VF_HIP(hipEventCreateWithFlags(&evHier, hipEventDisableTiming));
    }
    void wait_hier() {
        if (!hierPending) return;
        VF_HIP(hipStreamWaitEvent(st(), evHier, 0));
        hierPending = false;
    }
    // With useGraphs the factorization's ~90 small launches are replayed from a captured HIP graph: issued
    // one by one they keep the host busy for milliseconds while the main stream has nothing queued. The
    // graph is captured again when its operands (buffers, sizes, coarsest grid) change.
    GraphGuard factorGraph;
    std::string factorKey;
    // The factorization is issued from a worker thread: even replayed from a graph, its launches take the
    // issuing thread milliseconds while the device is busy, time in which the main stream would have
    // nothing queued. finish_factor joins the thread before the main stream waits for evFactor.
    std::thread factorThread;
    std::exception_ptr factorError;
    void join_factor() {
        if (factorThread.joinable()) factorThread.join();
        if (factorError) {
            std::exception_ptr e = factorError;
            factorError = nullptr;
            std::rethrow_exception(e);
        }
    }
    int bandKb = 1;
    DBuf<int> dRedToFull, dRed;
    int nred = 0, ldA = 0;
    bool stiffnessReady = false;
    uint64_t stiffnessVersion = 0;
    // Partial Galerkin updates after layers were removed from the top of the fabrication mask
    // (MultigridSolver.hh:907-1017): only the coarse element rows holding the newly masked fine layers are
    // recomputed, only the block rows of the nodes they touch reassembled; the coarsest level is refactored.
    // The recomputed rows come from the same kernels and inputs as in a full rebuild, so the operators are
    // identical to a full rebuild's.
    bool partialUpdates = true;
    uint64_t stiffBase = 0;
    int64_t stiffLayer = 0;    // first masked fine element layer at the last update
    int64_t partialCount = 0;  // partial updates made
    std::vector<std::array<int, 2>> dirtyRows[2];  // per precision and level: node rows [lo, hi] to reassemble
    // PCG work
    DBuf<double> pd, pAd, ps, pr, px, pb;
    // Block-diagonal coarsest level (a stacked layer-by-layer batch): segment k owns the coarsest node planes
    // [k Pc, k Pc + Wc]; the planes between segments border only void elements and are constrained, so the
    // reduced operator splits into one block per segment. Each block is factored and solved on its own, all
    // blocks in one batched launch sequence (set `seg` before the first stiffness update).
    struct SegCoarse {
        int nseg = 0, Pc = 0, Wc = 0;
        Grid gsub{};
        std::vector<int> n;
        std::vector<int> ydet;  // first detached coarsest row of each segment
        const int *state = nullptr;  // device: segments in state 2 (done) are left out of the coarse solve
        int nb = 1, kb = 1, ld = 64;
        DBuf<int> dn, dnb, red, redToFull, flags;
        DBuf<double> B, Linv, X, XT, tin, tout;
        bool pending = false;
    };
    std::unique_ptr<SegCoarse> seg;
    // the stacked batches of the last batched layer-by-layer evaluation on this solver: their grids, masks,
    // hierarchy buffers and work vectors are reused by the next evaluation with the same layers and
    // constraints (the moduli, hierarchy and factors are recomputed every time)
    std::vector<std::shared_ptr<StackedBatch>> stackCache;
    // the coarsened element matrices (Chat, the level-2 path matrices and the Walsh form of Chat_0) of
    // another solver, so that both build identical level operators from the same moduli
    void copy_coarsening(const vf_mg &o) {
        std::copy(o.Chat, o.Chat + 8 * 576, Chat);
        std::copy(o.phi, o.phi + 512, phi);
        auto dup = [&](auto &dst, const auto &src) {
            dst.resize(src.size());
            d2d(dst.data(), src.data(), src.size(), st());
        };
        dup(dChat, o.dChat); dup(dChat32, o.dChat32); dup(dphi, o.dphi); dup(dM2, o.dM2);
        dup(dCw, o.dCw); dup(dCw32, o.dCw32);
    }
    // level-0 views used by the cycles (set per preconditioner call)
    double *x0v = nullptr, *b0v = nullptr;
    // what debug_get_x/b (MultigridSolver.hh:1158-1159) read: the level buffers of the last cycle, in the
    // precision it ran in (-1: no cycle yet); level 0 is the last cycle's fine iterate and right-hand side
    int dbgPrec = -1;
    const void *dbgX0 = nullptr, *dbgB0 = nullptr;

    hipStream_t st() const { return sim->stream; }

    void sync_masks() {
        sim->refresh_grid();
        lv[0]->g = sim->g;
        for (int l = 1; l <= L; ++l) {
            Grid &g = lv[l]->g;
            if (std::isinf(sim->maskHeight)) {
                g.ydet = g.nn[1];
                g.ymask = g.ne[1];
            } else {
                const int64_t fm = (int64_t)std::ceil(sim->maskHeight / lv[l]->h[1] - 1e-10);
                g.ymask = (int)std::min<int64_t>(fm, g.ne[1]);
                g.ydet = (int)std::min<int64_t>(fm + 1, g.nn[1]);
            }
        }
    }

    ~vf_mg() {
        if (factorThread.joinable()) factorThread.join();
        if (fstream) {
            (void)hipStreamSynchronize(fstream);
            (void)hipStreamDestroy(fstream);
        }
        if (gstream) {
            (void)hipStreamSynchronize(gstream);
            (void)hipStreamDestroy(gstream);
        }
        if (evModuli) (void)hipEventDestroy(evModuli);
        if (evHier) (void)hipEventDestroy(evHier);
        if (evGalerkin) (void)hipEventDestroy(evGalerkin);
        if (evFactor) (void)hipEventDestroy(evFactor);
        for (hipEvent_t e : kdEv)
            if (e) (void)hipEventDestroy(e);
    }
    // measurement of the fine K d of PCG where it runs: HIP events on the solver's stream around each launch
    // (recorded into the replayed graph as well), summed after each iteration's host read of <r,r>
    bool timeKd = false;
    hipEvent_t kdEv[2] = {nullptr, nullptr};
    double kdMs = 0;
    int64_t kdCount = 0;
    vf_mg(const vf_mg &) = delete;
    vf_mg &operator=(const vf_mg &) = delete;

    vf_mg(vf_sim *s, std::unique_ptr<vf::planar::PMG> p) : sim(s), L(p->L), p2(std::move(p)) {}
    vf_mg(vf_sim *s, int levels) : sim(s), L(levels) {
        if (levels < 0) throw std::runtime_error("numCoarseningLevels must be >= 0");
        auto l0 = std::make_unique<LevelData>();
        l0->g = sim->g;
        std::copy(sim->h, sim->h + 3, l0->h);
        l0->hmask = sim->hmask;
        lv.push_back(std::move(l0));
        int64_t ne[3] = {sim->ne[0], sim->ne[1], sim->ne[2]};
        for (int l = 1; l <= L; ++l) {
            for (int d = 0; d < 3; ++d) {
                if (ne[d] % 2)
                    throw std::runtime_error(
                        "Grid size currently must be divisible by 2^numCoarseningLevels (nonuniform coarsening not yet "
                        "implemented)");
                ne[d] /= 2;
            }
            auto c = std::make_unique<LevelData>();
            c->g = make_grid(ne);
            for (int d = 0; d < 3; ++d) c->h[d] = (sim->hi[d] - sim->lo[d]) / ne[d];
            // Dirichlet coarsening (MultigridSolver.hh:58-103): along an axis where the constrained
            // fine node index is even the coarse node is i/2; where it is odd the fine node lies
            // inside a coarse edge/face/cell and both neighbours (i-1)/2, (i+1)/2 are constrained.
            const LevelData &f = *lv.back();
            c->hmask.assign(c->g.nnodes, 0);
            const int fnz = f.g.nn[2], fny = f.g.nn[1];
            for (int n = 0; n < f.g.nnodes; ++n) {
                const uint8_t m = f.hmask[n];
                if (!m) continue;
                const int id[3] = {n / (fny * fnz), (n / fnz) % fny, n % fnz};
                int a0[3], a1[3];
                for (int d = 0; d < 3; ++d) {
                    if (id[d] % 2 == 0) a0[d] = a1[d] = id[d] / 2;
                    else { a0[d] = (id[d] - 1) / 2; a1[d] = (id[d] + 1) / 2; }
                }
                for (int x = a0[0]; x <= a1[0]; ++x)
                    for (int y = a0[1]; y <= a1[1]; ++y)
                        for (int z = a0[2]; z <= a1[2]; ++z)
                            c->hmask[(x * c->g.nn[1] + y) * c->g.nn[2] + z] |= m;
            }
            upload_dirichlet(c->hmask, c->dmask, c->dnodes, c->dnmask, c->ndir, st());
            lv.push_back(std::move(c));
        }
        // coarsened fine K0s (MultigridSolver.hh:116-120) and the level-2 path matrices
        child_phi(phi);
        std::fill(Chat, Chat + 8 * 576, 0.0);
        for (int fi = 0; fi < 8; ++fi) coarsen_into(phi, fi, sim->K0h, Chat + fi * 576);
        dChat.resize(8 * 576);
        dChat32.resize(8 * 576);
        dphi.resize(512);
        std::vector<float> c32(Chat, Chat + 8 * 576);
        h2d(dChat.data(), Chat, 8 * 576, st());
        h2d(dChat32.data(), c32.data(), 8 * 576, st());
        // child 0's matrix in the Walsh basis for the element-wise level-1 product: Cw = W^T Chat_0 W / 64
        {
            auto wal = [](int l, int k) {
                int v = 1;
                for (int a = 0; a < 3; ++a)
                    if ((k >> (2 - a)) & 1) v *= ((l >> (2 - a)) & 1) ? 1 : -1;
                return v;
            };
            std::vector<double> cw(576, 0.0);
            for (int k = 0; k < 8; ++k)
                for (int a = 0; a < 3; ++a)
                    for (int kp = 0; kp < 8; ++kp)
                        for (int c = 0; c < 3; ++c) {
                            double acc = 0;
                            for (int l = 0; l < 8; ++l)
                                for (int m = 0; m < 8; ++m)
                                    acc += wal(l, k) * Chat[(3 * l + a) * 24 + 3 * m + c] * wal(m, kp);
                            cw[(3 * k + a) * 24 + 3 * kp + c] = acc / 64.0;
                        }
            std::vector<float> cw32(cw.begin(), cw.end());
            dCw.resize(576);
            dCw32.resize(576);
            h2d(dCw.data(), cw.data(), 576, st());
            h2d(dCw32.data(), cw32.data(), 576, st());
        }
        h2d(dphi.data(), phi, 512, st());
        if (L >= 2) {
            std::vector<double> M2(64 * 576, 0.0);
            for (int f2 = 0; f2 < 8; ++f2)
                for (int f1 = 0; f1 < 8; ++f1) coarsen_into(phi, f2, Chat + f1 * 576, &M2[(8 * f2 + f1) * 576]);
            dM2.resize(M2.size());
            h2d(dM2.data(), M2.data(), M2.size(), st());
        }
        sync(st());
        sync_masks();
    }

    // ---------------------------------------------------------------- buffers
    template <typename T>
    T *X(int l);
    template <typename T>
    T *B(int l);
    template <typename T>
    T *R(int l);
    template <typename T>
    void ensure_buffers(int from) {
        for (int l = from; l <= L; ++l) {
            LevelData &d = *lv[l];
            const size_t n = 3 * (size_t)d.g.nnodes;
            if constexpr (std::is_same<T, double>::value) {
                d.x.resize(n); d.b.resize(n); d.r.resize(n);
            } else {
                d.x32.resize(n); d.b32.resize(n); d.r32.resize(n);
            }
        }
    }

    const uint8_t *dm(int l) const { return l == 0 ? sim->dmask.data() : lv[l]->dmask.data(); }
    template <typename T>
    void zero_dirichlet(int l, T *u) {
        if (l == 0) launch_zero_dirichlet<T>(sim->ndir, sim->dnodes.data(), sim->dnmask.data(), sim->g.nnodes, u, st());
        else launch_zero_dirichlet<T>(lv[l]->ndir, lv[l]->dnodes.data(), lv[l]->dnmask.data(), lv[l]->g.nnodes, u, st());
    }

    // ---------------------------------------------------------------- level operators
    template <typename T>
    const T *K0p() const;
    template <typename T>
    const T *Ep() const;
    template <typename T>
    const T *Chatp() const;
    template <typename T>
    const T *Cwp() const;
    template <typename T>
    const T *blocksp(int l) const;

    template <typename T>
    const T *Fp() const;
    template <typename T>
    const T *Wp() const;
    // Fine-level kernel family per precision (index 0 FP64, 1 FP32). K u: 0 register-gather, 1 LDS-tiled
    // node rows, 2 element-centric Walsh; smoother: 0 register-gather (one launch per colour), 1 LDS-tiled
    // two-launch sweep. -1 = not tuned yet. See tune().
    int kuChoice[2] = {-1, -1}, gsChoice[2] = {-1, -1};
    template <typename T>
    static constexpr int pidx() { return std::is_same<T, double>::value ? 0 : 1; }
    template <typename T>
    int ku_family() const {
        switch (g_kernel_variant) {
        case 1: return 0;
        case 2: return sim->structured ? 1 : 0;
        case 3:
        case 4: return sim->walsh ? 2 : (sim->structured ? 1 : 0);
        default: {
            const int c = kuChoice[pidx<T>()];
            if (c >= 0) return c;
            return sim->walsh ? 2 : (sim->structured ? 1 : 0);
        }
        }
    }
    // smoother family: 0 register gather (dense K0), 1 LDS-tiled two-launch sweep, 2 register gather on
    // the structured element matrix
    template <typename T>
    int gs_family() const {
        if (!sim->structured || g_kernel_variant == 1) return 0;
        if (g_kernel_variant == 4) return 2;
        if (g_kernel_variant >= 2) return 1;
        const int c = gsChoice[pidx<T>()];
        return c >= 0 ? c : 2;
    }
    template <typename T>
    bool tiled_gs() const { return gs_family<T>() == 1; }
    // Autotuner: time each applicable kernel family on this grid (HIP events on the solver's stream, best of
    // three runs after a warm-up, with nothing else running on the device) and keep the fastest per operation
    // and precision.
    template <typename T>
    void tune() {
        ensure_blocks<T>(true);
        const int p = pidx<T>();
        if (g_kernel_variant != 0 || !sim->structured || (kuChoice[p] >= 0 && gsChoice[p] >= 0)) return;
        wait_hier();
        // nothing else may run meanwhile: the coarsest factorization (its own stream) would compete for the CUs
        finish_factor();
        sync(st());
        const Grid &g = lv[0]->g;
        const size_t n = 3 * (size_t)g.nnodes;
        hipStream_t s = st();
        DBuf<T> u(n), y(n), rhs(n);
        launch_fill<T>(n, u.data(), T(1e-3), s);
        launch_fill<T>(n, rhs.data(), T(1), s);
        hipEvent_t a = nullptr, b = nullptr;
        struct EventGuard {
            hipEvent_t *e0, *e1;
            ~EventGuard() {
                if (*e0) (void)hipEventDestroy(*e0);
                if (*e1) (void)hipEventDestroy(*e1);
            }
        } guard{&a, &b};  // the events live until every timing below is done
        VF_HIP(hipEventCreate(&a));
        VF_HIP(hipEventCreate(&b));
        auto best = [&](auto &&fn) {
            float ms = 1e30f;
            fn();
            for (int r = 0; r < 3; ++r) {
                VF_HIP(hipEventRecord(a, s));
                fn();
                VF_HIP(hipEventRecord(b, s));
                VF_HIP(hipEventSynchronize(b));
                float t = 0;
                VF_HIP(hipEventElapsedTime(&t, a, b));
                ms = std::min(ms, t);
            }
            return ms;
        };
        const uint8_t *mask = dm(0);
        float ku[3] = {1e30f, 1e30f, 1e30f};
        ku[0] = best([&] { launch_fine_apply<T>(g, K0p<T>(), Ep<T>(), u.data(), y.data(), rhs.data(), mask,
                                                AM_RESIDUAL, s); });
        ku[1] = best([&] { launch_fine_apply_tiled<T>(g, Fp<T>(), Ep<T>(), u.data(), y.data(), rhs.data(), mask,
                                                      AM_RESIDUAL, s); });
        if (sim->walsh)
            ku[2] = best([&] { launch_fine_apply_walsh<T>(g, Wp<T>(), Ep<T>(), u.data(), y.data(), rhs.data(), mask,
                                                          AM_RESIDUAL, s); });
        const float gsT = best([&] {
            launch_fine_gs_phase<T>(g, Fp<T>(), Ep<T>(), u.data(), u.data(), y.data(), rhs.data(), mask, 0, 1, s);
            launch_fine_gs_phase<T>(g, Fp<T>(), Ep<T>(), u.data(), y.data(), y.data(), rhs.data(), mask, 1, 1, s);
        });
        const float gsG = best([&] {
            for (int c = 0; c < 8; ++c) launch_fine_gs<T>(g, K0p<T>(), Ep<T>(), y.data(), rhs.data(), mask, c, 1, s);
        });
        const float gsF = best([&] {
            for (int c = 0; c < 8; ++c)
                launch_fine_gs_gather_f<T>(g, Fp<T>(), Ep<T>(), y.data(), rhs.data(), mask, c, 1, s);
        });
        int kbest = 0;
        for (int i = 1; i < 3; ++i)
            if (ku[i] < ku[kbest]) kbest = i;
        kuChoice[p] = kbest;
        gsChoice[p] = gsT <= gsG ? (gsT <= gsF ? 1 : 2) : (gsG <= gsF ? 0 : 2);
        // level 1: residual and smoother each from the assembled block rows, on the fly on the vector units
        // or on the fly on the matrix cores, whichever is fastest here
        l1Tuned[p][0] = l1Tuned[p][1] = -1;
        if (l1Blocks && L >= 2) {
            const Grid &g1 = lv[1]->g;
            float tr[3], ts[3];
            tr[0] = best([&] { launch_block_apply<T>(g1, blocksp<T>(1), u.data(), y.data(), rhs.data(), dm(1),
                                                     AM_RESIDUAL, s); });
            tr[1] = best([&] { launch_l1_apply<T>(g1, g, Chatp<T>(), Cwp<T>(), Ep<T>(), u.data(), y.data(), rhs.data(), dm(1),
                                                  AM_RESIDUAL, sim->structured, s); });
            tr[2] = best([&] { launch_l1_apply_mfma<T>(g1, g, Chatp<T>(), Ep<T>(), u.data(), y.data(), rhs.data(),
                                                       dm(1), AM_RESIDUAL, s); });
            // one forward sweep, from the same start each time
            const size_t n1 = 3 * (size_t)g1.nnodes;
            DBuf<T> u1(n1), b1(n1);
            launch_fill<T>(n1, b1.data(), T(1), s);
            auto sweep = [&](int form) {
                launch_fill<T>(n1, u1.data(), T(1e-3), s);
                for (int c = 0; c < 8; ++c) {
                    if (form == 0) launch_block_gs<T>(g1, blocksp<T>(1), u1.data(), b1.data(), dm(1), c, 1, s);
                    else if (form == 1)
                        launch_l1_gs<T>(g1, g, Chatp<T>(), Ep<T>(), u1.data(), b1.data(), dm(1), c, 1, sim->structured, s);
                    else launch_l1_gs_mfma<T>(g1, g, Chatp<T>(), Ep<T>(), u1.data(), b1.data(), dm(1), c, 1, s);
                }
            };
            for (int f = 0; f < 3; ++f) ts[f] = best([&] { sweep(f); });
            l1Tuned[p][0] = tr[0] <= tr[1] ? (tr[0] <= tr[2] ? 0 : 2) : (tr[1] <= tr[2] ? 1 : 2);
            l1Tuned[p][1] = ts[0] <= ts[1] ? (ts[0] <= ts[2] ? 0 : 2) : (ts[1] <= ts[2] ? 1 : 2);
            for (int f = 0; f < 3; ++f) {
                tuneMs[p][6 + f] = tr[f];
                tuneMs[p][9 + f] = ts[f];
            }
        }
        tuneMs[p][0] = ku[0];
        tuneMs[p][1] = ku[1];
        tuneMs[p][2] = sim->walsh ? ku[2] : 0.0f;
        tuneMs[p][3] = gsG;
        tuneMs[p][4] = gsT;
        tuneMs[p][5] = gsF;
    }
    float tuneMs[2][12] = {};
    // level-1 form of the residual ([p][0]) and of the smoother ([p][1]) the tuner chose: 0 assembled block
    // rows, 1 on the fly on the vector units, 2 on the fly on the matrix cores; -1 not tuned
    int l1Tuned[2][2] = {{-1, -1}, {-1, -1}};
    int l1_form(int p, int op) const {
        if (!(l1Blocks && L >= 2)) return g_kernel_variant == 5 ? 2 : 1;  // no block rows: on the fly
        return l1Tuned[p][op] >= 0 ? l1Tuned[p][op] : 0;
    }
    // level-1 block rows are read by nothing once the tuner put both level-1 operations on the fly
    bool l1_blocks_unused(int p) const { return l1Tuned[p][0] > 0 && l1Tuned[p][1] > 0; }

    // Ad = K d on the fine level (Dirichlet rows zero) and *dAd = <d, Ad> over the active nodes, fused into
    // the Walsh-basis product when that family is in use
    void apply_dot(const double *d, double *Ad, double *partials, double *dAd);
    template <typename T>
    void apply(int l, const T *u, T *y, const T *b, int mode, bool useMask) {
        const uint8_t *mask = useMask ? dm(l) : nullptr;
        const int fam = l == 0 && (mode == AM_SET || mode == AM_RESIDUAL) ? ku_family<T>() : 0;
        if (l == 0 && fam == 2) launch_fine_apply_walsh<T>(lv[0]->g, Wp<T>(), Ep<T>(), u, y, b, mask, mode, st());
        else if (l == 0 && fam == 1) launch_fine_apply_tiled<T>(lv[0]->g, Fp<T>(), Ep<T>(), u, y, b, mask, mode, st());
        else if (l == 0) launch_fine_apply<T>(lv[0]->g, K0p<T>(), Ep<T>(), u, y, b, mask, mode, st());
        else if (l == 1 && l1_form(pidx<T>(), 0) == 1)
            launch_l1_apply<T>(lv[1]->g, lv[0]->g, Chatp<T>(), Cwp<T>(), Ep<T>(), u, y, b, mask, mode, sim->structured,
                               st());
        else if (l == 1 && l1_form(pidx<T>(), 0) == 2)
            launch_l1_apply_mfma<T>(lv[1]->g, lv[0]->g, Chatp<T>(), Ep<T>(), u, y, b, mask, mode, st());
        else if (l < L) launch_block_apply<T>(lv[l]->g, blocksp<T>(l), u, y, b, mask, mode, st());
        else {
            if constexpr (std::is_same<T, double>::value) {
                if (mode == AM_RESIDUAL) {
                    d2d(y, b, 3 * (size_t)lv[l]->g.nnodes, st());
                    launch_elem_apply<double>(lv[l]->g, lv[l]->Ke.data(), u, y, AM_SUB, st());
                } else {
                    launch_elem_apply<double>(lv[l]->g, lv[l]->Ke.data(), u, y, mode, st());
                }
                if (useMask) zero_dirichlet<double>(l, y);
            } else {
                throw std::runtime_error("coarsest-level operator is FP64 only");
            }
        }
    }
    // one sweep from uin into uout (uout != uin) on levels with a two-launch smoother; zeroIn: uin is zero
    // (its memory is not read, only its detached rows need to hold zeros)
    template <typename T>
    void sweep_to(int l, const T *uin, T *uout, const T *b, bool forward, bool zeroIn = false) {
        const int p0 = forward ? 0 : 1;
        launch_fine_gs_phase<T>(lv[0]->g, Fp<T>(), Ep<T>(), uin, uin, uout, b, dm(0), p0, forward, st(),
                                zeroIn ? 3 : 0);
        launch_fine_gs_phase<T>(lv[0]->g, Fp<T>(), Ep<T>(), uin, uout, uout, b, dm(0), 1 - p0, forward, st(),
                                zeroIn ? 1 : 0);
    }
    template <typename T>
    bool paired(int l) const { return l == 0 && tiled_gs<T>(); }
    template <typename T>
    void smooth(int l, T *u, const T *b, bool forward) {
        if (paired<T>(l)) {
            T *w = R<T>(l);
            sweep_to<T>(l, u, w, b, forward);
            d2d(u, w, 3 * (size_t)lv[l]->g.nnodes, st());
            return;
        }
        for (int t = 0; t < 8; ++t) {
            const int color = forward ? t : 7 - t;
            if (l == 0 && gs_family<T>() == 2)
                launch_fine_gs_gather_f<T>(lv[0]->g, Fp<T>(), Ep<T>(), u, b, dm(0), color, forward, st());
            else if (l == 0) launch_fine_gs<T>(lv[0]->g, K0p<T>(), Ep<T>(), u, b, dm(0), color, forward, st());
            else if (l == 1 && l1_form(pidx<T>(), 1) == 1)
                launch_l1_gs<T>(lv[1]->g, lv[0]->g, Chatp<T>(), Ep<T>(), u, b, dm(1), color, forward, sim->structured,
                                st());
            else if (l == 1 && l1_form(pidx<T>(), 1) == 2)
                launch_l1_gs_mfma<T>(lv[1]->g, lv[0]->g, Chatp<T>(), Ep<T>(), u, b, dm(1), color, forward, st());
            else launch_block_gs<T>(lv[l]->g, blocksp<T>(l), u, b, dm(l), color, forward, st());
        }
    }
    template <typename T>
    void coarse_solve(const T *b, T *x) {
        if (seg) {
            ensure_factor();
            seg_finish();
            const SegCoarse &S = *seg;
            const Grid &g = lv[L]->g;
            launch_coarse_chol_apply_batched<T>(S.nseg, S.dn.data(), S.ld, S.X.data(), S.XT.data(), S.redToFull.data(),
                                                (size_t)S.ld, g.nnodes, (size_t)S.Pc * g.nn[1] * g.nn[2], b, x,
                                                S.tin.data(), S.tout.data(), st(), S.state);
            return;
        }
        ensure_factor();
        finish_factor();
        launch_coarse_chol_apply<T>(nred, Ainv.data(), AinvT.data(), dRedToFull.data(), lv[L]->g.nnodes, b, x, ctmp_in.data(),
                                    ctmp_out.data(), st());
    }

    // ---------------------------------------------------------------- cycles (MultigridSolver.hh:587-658)
    // enforceDirichletConditions(l, x, residualSystem) (MultigridSolver.hh:516-524): coarse levels
    // always solve for corrections with zero Dirichlet values.
    template <typename T>
    void dirichlet_for_cycle(int l, T *x, bool residual) {
        if (l == 0 && !residual) {
            if constexpr (std::is_same<T, double>::value)
                launch_enforce_dirichlet(sim->ndir, sim->dnodes.data(), sim->dnmask.data(), sim->g.nnodes,
                                         sim->dvals.data(), x, st());
            else throw std::runtime_error("non-residual cycles run in FP64");
        } else {
            zero_dirichlet<T>(l, x);
        }
    }
    template <typename T>
    // xZero: the caller left x at zero (the restriction zeroes the coarse iterate), so its Dirichlet
    // components need no clearing
    void vcycle(int l, int nu, bool residual = true, bool xZero = false) {
        if (l >= 2 || (l == 1 && l1Blocks && !l1_blocks_unused(pidx<T>()))) wait_hier();
        if (l == L) { coarse_solve<T>(B<T>(l), X<T>(l)); return; }
        T *x = X<T>(l);
        const T *b = B<T>(l);
        if (!xZero) dirichlet_for_cycle<T>(l, x, residual);
        if (paired<T>(l)) {
            // sweeps alternate between x and the residual buffer; the residual goes to whichever of
            // the two does not hold the iterate, and 2 nu sweeps bring the iterate back to x
            T *cur = x, *alt = R<T>(l);
            for (int i = 0; i < nu; ++i) { sweep_to<T>(l, cur, alt, b, true, xZero && i == 0); std::swap(cur, alt); }
            apply<T>(l, cur, alt, b, AM_RESIDUAL, true);
            launch_restrict<T>(lv[l]->g, lv[l + 1]->g, alt, B<T>(l + 1), st(), X<T>(l + 1));
            vcycle<T>(l + 1, nu, true, true);
            launch_interpolate<T>(lv[l]->g, lv[l + 1]->g, X<T>(l + 1), cur, 1, st());
            for (int i = 0; i < nu; ++i) { sweep_to<T>(l, cur, alt, b, !symmetricGS); std::swap(cur, alt); }
            if (cur != x) d2d(x, cur, 3 * (size_t)lv[l]->g.nnodes, st());
            return;
        }
        for (int i = 0; i < nu; ++i) smooth<T>(l, x, b, true);
        apply<T>(l, x, R<T>(l), b, AM_RESIDUAL, true);
        launch_restrict<T>(lv[l]->g, lv[l + 1]->g, R<T>(l), B<T>(l + 1), st(), X<T>(l + 1));
        vcycle<T>(l + 1, nu, true, true);
        launch_interpolate<T>(lv[l]->g, lv[l + 1]->g, X<T>(l + 1), x, 1, st());
        for (int i = 0; i < nu; ++i) smooth<T>(l, x, b, !symmetricGS);
    }
    template <typename T>
    void fmg(int l, int nu, bool residual = true) {
        if (l == L) { wait_hier(); coarse_solve<T>(B<T>(l), X<T>(l)); return; }
        launch_restrict<T>(lv[l]->g, lv[l + 1]->g, B<T>(l), B<T>(l + 1), st());
        fmg<T>(l + 1, nu, residual);
        launch_interpolate<T>(lv[l]->g, lv[l + 1]->g, X<T>(l + 1), X<T>(l), 0, st());
        vcycle<T>(l, nu, residual);
    }

    // MultigridSolver::solve (MultigridSolver.hh:546-575) in FP64 on device vectors
    void solve(double *x, const double *f, int numSteps, int nu, bool stiffnessUpdated, bool zeroDirichlet,
               bool fullMG, vf_mg_callback cb, void *user) {
        sync_masks();
        if (!stiffnessUpdated) update_stiffness();
        else ensure_stiffness();
        if (numSteps == 0) return;
        const size_t n0 = 3 * (size_t)lv[0]->g.nnodes;
        std::vector<double> hx;
        if (L == 0) {
            coarse_solve<double>(f, x);
            if (cb) {
                hx.resize(n0); d2h(hx.data(), x, n0, st()); sync(st());
                if (cb(user, 0, hx.data())) throw std::runtime_error("solve aborted by callback");
            }
            return;
        }
        ensure_buffers<double>(1);
        tune<double>();
        b0v = const_cast<double *>(f);
        x0v = x;
        for (int i = 0; i < numSteps; ++i) {
            if (fullMG && i == 0) fmg<double>(0, nu, zeroDirichlet);
            else vcycle<double>(0, nu, zeroDirichlet);
            if (cb) {
                hx.resize(n0); d2h(hx.data(), x, n0, st()); sync(st());
                if (cb(user, i, hx.data())) { b0v = x0v = nullptr; throw std::runtime_error("solve aborted by callback"); }
            }
        }
        dbgPrec = VF_PREC_FP64;
        dbgX0 = x;
        dbgB0 = f;
        b0v = x0v = nullptr;
    }

    // applyPreconditionerInv (MultigridSolver.hh:577-580) on device vectors (SoA, fine level)
    // resident (FP32 cycles only): the right-hand side is already in lv[0]->b32 and the result stays in
    // lv[0]->x32 (r and s are not touched)
    // the fine iterate of the first cycle starts at zero: with the two-launch smoother its first sweep
    // reads it as zero without loading it, so only the detached rows (which no sweep writes) are cleared
    template <typename T>
    void zero_start(T *x, bool fullMG) {
        if (paired<T>(0) && !fullMG) launch_zero_detached<T>(lv[0]->g, x, st());
        else launch_fill<T>(3 * (size_t)lv[0]->g.nnodes, x, T(0), st());
    }
    void precondition(const double *r, double *s, int mgIter, int nu, bool fullMG, int prec, bool resident = false) {
        sync_masks();
        const size_t n0 = 3 * (size_t)lv[0]->g.nnodes;
        if (nu == 0 || mgIter == 0) { d2d(s, r, n0, st()); return; }
        if (L == 0) { coarse_solve<double>(r, s); return; }
        ensure_stiffness();
        if (prec == VF_PREC_FP64) {
            ensure_buffers<double>(1);
            tune<double>();
            b0v = const_cast<double *>(r);
            x0v = s;
            zero_start<double>(s, fullMG);
            for (int i = 0; i < mgIter; ++i) {
                if (fullMG && i == 0) fmg<double>(0, nu);
                else vcycle<double>(0, nu, true, i == 0);
            }
            dbgPrec = VF_PREC_FP64;
            dbgX0 = s;
            dbgB0 = r;
        } else {
            ensure_buffers<float>(0);
            tune<float>();
            if (!resident) launch_convert<float, double>(n0, r, lv[0]->b32.data(), st());
            zero_start<float>(lv[0]->x32.data(), fullMG);
            for (int i = 0; i < mgIter; ++i) {
                if (fullMG && i == 0) fmg<float>(0, nu);
                else vcycle<float>(0, nu, true, i == 0);
            }
            dbgPrec = VF_PREC_FP32;
            dbgX0 = lv[0]->x32.data();
            dbgB0 = lv[0]->b32.data();
            if (!resident) launch_convert<double, float>(n0, lv[0]->x32.data(), s, st());
        }
        b0v = x0v = nullptr;
    }

    // the level-0 vectors of an API-level cycle live in caller or staging memory: keep a copy for debug_get
    void dbg_keep_level0(size_t n0) {
        if (dbgPrec != VF_PREC_FP64 || (dbgX0 == lv[0]->x.data() && dbgB0 == lv[0]->b.data())) return;
        lv[0]->x.resize(n0);
        lv[0]->b.resize(n0);
        d2d(lv[0]->x.data(), (const double *)dbgX0, n0, st());
        d2d(lv[0]->b.data(), (const double *)dbgB0, n0, st());
        dbgX0 = lv[0]->x.data();
        dbgB0 = lv[0]->b.data();
    }

    // ---------------------------------------------------------------- Galerkin setup
    void ensure_stiffness() {
        if (stiffnessReady && stiffnessVersion == sim->version) return;
        const int64_t layer = std::min<int64_t>(sim->firstMasked, sim->ne[1]);
        if (partialUpdates && stiffnessReady && L >= 2 && sim->baseVersion == stiffBase && layer < stiffLayer)
            partial_update(layer, stiffLayer);
        else
            update_stiffness();
    }
    // the fine element layers [m, c) were masked since the last update
    void partial_update(int64_t m, int64_t c) {
        sync_masks();
        hipStream_t s = st();
        wait_hier();
        prepare_coarsest();
        // element rows of level l that hold fine layers of [m, c) (the coarsened bands, MultigridSolver.hh:913-922)
        std::vector<std::array<int, 2>> band(L + 1);
        int64_t b0 = m, b1 = c;
        band[0] = {(int)b0, (int)b1};
        for (int l = 1; l <= L; ++l) {
            b0 /= 2;
            b1 = (b1 + 1) / 2;
            band[l] = {(int)b0, (int)b1};
        }
        launch_galerkin_l2_elems(lv[2]->g, lv[0]->g, dM2.data(), sim->E.data(), lv[2]->Ke.data(), s, band[2][0],
                                 band[2][1]);
        for (int l = 3; l <= L; ++l)
            launch_galerkin_recurse(lv[l]->g, lv[l - 1]->g, dphi.data(), lv[l - 1]->Ke.data(), lv[l]->Ke.data(), s,
                                    band[l][0], band[l][1]);
        // the block rows of the nodes on those element rows are reassembled by the next cycle that uses them
        for (int p = 0; p < 2; ++p) {
            if (!blocksReady[p]) continue;
            auto &d = dirtyRows[p];
            if ((int)d.size() < L + 1) d.assign(L + 1, {1, 0});
            for (int l = 1; l < L; ++l) {
                const int lo = band[l][0], hi = band[l][1];
                if (d[l][0] > d[l][1]) d[l] = {lo, hi};
                else d[l] = {std::min(d[l][0], lo), std::max(d[l][1], hi)};
            }
        }
        launch_factor(s);
        stiffnessReady = true;
        stiffnessVersion = sim->version;
        stiffLayer = m;
        ++partialCount;
    }
    // updateStiffnessMatrices (MultigridSolver.hh:846-905)
    void update_stiffness() {
        sync_masks();
        hipStream_t s = st();
        prepare_coarsest();  // host index maps and their upload, before the device work is queued
        if (L == 0) {
            LevelData &c = *lv[0];
            c.Ke.resize((size_t)c.g.nelems * 576);
            // Ke = E_e K0: reuse the level-1 kernel shape via a scaled copy on host-sized grids
            std::vector<double> E(c.g.nelems), Ke((size_t)c.g.nelems * 576);
            d2h(E.data(), sim->E.data(), E.size(), s);
            sync(s);
            for (int e = 0; e < c.g.nelems; ++e)
                for (int k = 0; k < 576; ++k) Ke[(size_t)e * 576 + k] = E[e] * sim->K0h[k];
            h2d(c.Ke.data(), Ke.data(), Ke.size(), s);
        } else if (L == 1) {
            LevelData &c = *lv[1];
            c.Ke.resize((size_t)c.g.nelems * 576);
            launch_galerkin_l1_elems(c.g, lv[0]->g, dChat.data(), sim->E.data(), c.Ke.data(), s);
        } else {
            // level 1 as assembled block rows (HBM holds them easily; the reference applies this level
            // on the fly to save memory), unless the register-gather family is forced
            l1Blocks = g_kernel_variant != 1 && g_kernel_variant != 5;
            if (!l1Blocks) {
                lv[1]->blocks.release();
                lv[1]->blocks32.release();
            }
            // the buffers are (re)allocated before the chain is queued on the side stream
            for (int l = 2; l <= L; ++l) lv[l]->Ke.resize((size_t)lv[l]->g.nelems * 576);
            hipStream_t hs = s;
            if (overlapSetup) {
                ensure_gstream();
                VF_HIP(hipEventRecord(evModuli, s));
                VF_HIP(hipStreamWaitEvent(gstream, evModuli, 0));
                hs = gstream;
            }
            LevelData &c2 = *lv[2];
            launch_galerkin_l2_elems(c2.g, lv[0]->g, dM2.data(), sim->E.data(), c2.Ke.data(), hs);
            for (int l = 3; l <= L; ++l) {
                LevelData &c = *lv[l];
                launch_galerkin_recurse(c.g, lv[l - 1]->g, dphi.data(), lv[l - 1]->Ke.data(), c.Ke.data(), hs);
            }
            if (overlapSetup) {
                VF_HIP(hipEventRecord(evHier, gstream));
                hierPending = true;
            }
        }
        blocksReady[0] = blocksReady[1] = false;
        dirtyRows[0].clear();
        dirtyRows[1].clear();
        launch_factor(hierPending ? gstream : s);
        stiffnessReady = true;
        stiffnessVersion = sim->version;
        stiffBase = sim->baseVersion;
        stiffLayer = std::min<int64_t>(sim->firstMasked, sim->ne[1]);
    }
    template <typename T>
    DBuf<T> &block_buf(int l);
    // block rows of levels 1 (when assembled) .. L-1 in the precision T of the cycle about to use them,
    // summed in FP64 from the current element matrices
    // lazy: leave the solver's stream unordered with a pending hierarchy (the cycle waits where it first needs
    // level >= 2 data); otherwise the stream is ordered after the hierarchy on return
    template <typename T>
    void ensure_blocks(bool lazy = false) {
        ensure_stiffness();
        const int p = pidx<T>();
        // with the hierarchy still pending the assembly follows the Galerkin chain on its stream
        hipStream_t s = hierPending ? gstream : st();
        if (blocksReady[p]) {
            // after partial updates: only the stale node rows
            auto &d = dirtyRows[p];
            for (int l = 1; l < (int)d.size() && l < L; ++l) {
                if (d[l][0] > d[l][1]) continue;
                if (l == 1) {
                    if (l1Blocks && !l1_blocks_unused(p) && block_buf<T>(1).size())
                        launch_assemble_l1_blocks<T>(lv[1]->g, lv[0]->g, dChat.data(), sim->E.data(), block_buf<T>(1).data(),
                                                     sim->structured, s, d[l][0], d[l][1]);
                } else {
                    launch_assemble_blocks<T>(lv[l]->g, lv[l]->Ke.data(), block_buf<T>(l).data(), s, d[l][0], d[l][1]);
                }
            }
            d.clear();
            if (!lazy) wait_hier();
            return;
        }
        if (L >= 2 && l1Blocks && !l1_blocks_unused(p)) {
            LevelData &c1 = *lv[1];
            DBuf<T> &buf = block_buf<T>(1);
            buf.resize((size_t)c1.g.nnodes * 243);
            launch_assemble_l1_blocks<T>(c1.g, lv[0]->g, dChat.data(), sim->E.data(), buf.data(), sim->structured, s);
        }
        for (int l = 2; l < L; ++l) {
            LevelData &c = *lv[l];
            DBuf<T> &buf = block_buf<T>(l);
            buf.resize((size_t)c.g.nnodes * 243);
            launch_assemble_blocks<T>(c.g, c.Ke.data(), buf.data(), s);
        }
        blocksReady[p] = true;
        dirtyRows[p].clear();
        if (hierPending) VF_HIP(hipEventRecord(evHier, gstream));
        if (!lazy) wait_hier();
    }
    // the main stream waits for the coarsest factorization launched by the last update; a non-positive
    // pivot is reported here
    // A failure leaves the hierarchy marked stale, so every later solve rebuilds and refactors (and reports
    // the failure again) instead of running with the failed factor.
    void finish_factor() {
        if (!factorPending) return;
        factorPending = false;
        try {
            join_factor();
        } catch (...) {
            stiffnessReady = false;
            throw;
        }
        VF_HIP(hipStreamWaitEvent(st(), evFactor, 0));
        int flag = 0;
        d2h(&flag, cholFlag.data(), 1, fstream);
        sync(fstream);
        if (flag) {
            stiffnessReady = false;
            throw std::runtime_error("coarsest-level operator is not positive definite");
        }
    }
    void ensure_factor() {
        if (!stiffnessReady) update_stiffness();
    }
    // reduced coarsest operator: Dirichlet components and detached nodes removed
    // (findFixedVars, TensorProductSimulator.hh:1181-1195), then its FP64 inverse.
    // per-segment reduced maps (the same DoF order as a single structure's prepare_coarsest) and buffers
    void seg_prepare() {
        SegCoarse &S = *seg;
        const Grid &g = lv[L]->g;
        const std::vector<uint8_t> &mask = L == 0 ? sim->hmask : lv[L]->hmask;
        const int64_t nes[3] = {S.Wc, g.ne[1], g.ne[2]};
        S.gsub = make_grid(nes);
        const Grid &gs = S.gsub;
        const size_t rs = 3 * (size_t)gs.nnodes;
        std::vector<int> red(rs * S.nseg, -1);
        std::vector<std::vector<int>> r2f(S.nseg);
        S.n.assign(S.nseg, 0);
        int bw = 0, nmax = 0;
        for (int k = 0; k < S.nseg; ++k) {
            int *rk = red.data() + k * rs;
            for (int n = 0; n < gs.nnodes; ++n) {
                const int lx = n / (gs.nn[1] * gs.nn[2]), iy = (n / gs.nn[2]) % gs.nn[1], iz = n % gs.nn[2];
                const int m = ((k * S.Pc + lx) * g.nn[1] + iy) * g.nn[2] + iz;
                for (int a = 0; a < 3; ++a) {
                    const bool fixed = ((mask[m] >> a) & 1) || iy >= S.ydet[k];
                    if (!fixed) { rk[3 * n + a] = (int)r2f[k].size(); r2f[k].push_back(3 * n + a); }
                }
            }
            S.n[k] = (int)r2f[k].size();
            nmax = std::max(nmax, S.n[k]);
            for (int n = 0; n < gs.nnodes; ++n) {
                const int ix = n / (gs.nn[1] * gs.nn[2]), iy = (n / gs.nn[2]) % gs.nn[1], iz = n % gs.nn[2];
                for (int a = 0; a < 3; ++a) {
                    const int rr = rk[3 * n + a];
                    if (rr < 0) continue;
                    for (int dx = -1; dx <= 1; ++dx)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dz = -1; dz <= 1; ++dz) {
                                const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                                if (jx < 0 || jy < 0 || jz < 0 || jx >= gs.nn[0] || jy >= gs.nn[1] || jz >= gs.nn[2])
                                    continue;
                                const int mm = (jx * gs.nn[1] + jy) * gs.nn[2] + jz;
                                for (int c2 = 0; c2 < 3; ++c2) {
                                    const int cc = rk[3 * mm + c2];
                                    if (cc >= 0 && cc < rr) bw = std::max(bw, rr - cc);
                                }
                            }
                }
            }
        }
        S.kb = std::max(1, (bw + 63) / 64);
        S.nb = std::max(1, (nmax + 63) / 64);
        S.ld = S.nb * 64;
        const double bytes = (double)S.nseg * (2.0 * S.ld * S.ld + (double)band_storage(S.ld, S.kb)) * 8.0;
        if (S.ld > 24576 || bytes > 32e9)
            throw std::runtime_error("stacked coarsest level: " + std::to_string(nmax) +
                                     " unknowns per segment exceed the coarse-solver limit; use more coarsening levels");
        std::vector<int> r2fAll((size_t)S.nseg * S.ld, 0);
        for (int k = 0; k < S.nseg; ++k) std::copy(r2f[k].begin(), r2f[k].end(), r2fAll.begin() + (size_t)k * S.ld);
        hipStream_t s = st();
        S.red.resize(red.size());
        S.redToFull.resize(r2fAll.size());
        S.dn.resize(S.nseg);
        h2d(S.red.data(), red.data(), red.size(), s);
        h2d(S.redToFull.data(), r2fAll.data(), r2fAll.size(), s);
        h2d(S.dn.data(), S.n.data(), S.n.size(), s);
        std::vector<int> nbz(S.nseg);
        for (int k = 0; k < S.nseg; ++k) nbz[k] = std::max(1, (S.n[k] + 63) / 64);
        S.dnb.resize(S.nseg);
        h2d(S.dnb.data(), nbz.data(), nbz.size(), s);
        S.B.resize((size_t)S.nseg * band_storage(S.ld, S.kb));
        S.Linv.resize((size_t)S.nseg * S.nb * 4096);
        S.X.resize((size_t)S.nseg * S.ld * S.ld);
        S.XT.resize((size_t)S.nseg * S.ld * S.ld);
        S.tin.resize((size_t)S.nseg * S.ld);
        S.tout.resize((size_t)S.nseg * S.ld);
        S.flags.resize(S.nseg);
        sync(s);
    }
    void seg_factor(hipStream_t s) {
        SegCoarse &S = *seg;
        const Grid &g = lv[L]->g;
        band_cholesky_inverse_batched(S.gsub, lv[L]->Ke.data(), (size_t)S.Pc * g.ne[1] * g.ne[2] * 576, S.red.data(),
                                      3 * (size_t)S.gsub.nnodes, S.nseg, S.dn.data(), S.dnb.data(), S.nb, S.kb,
                                      S.B.data(),
                                      S.Linv.data(), S.X.data(), S.XT.data(), S.flags.data(), s);
        S.pending = true;
    }
    // a non-positive pivot in any block is reported at the first coarse solve after the factorization
    void seg_finish() {
        SegCoarse &S = *seg;
        if (!S.pending) return;
        S.pending = false;
        std::vector<int> f(S.nseg);
        d2h(f.data(), S.flags.data(), f.size(), st());
        sync(st());
        for (int k = 0; k < S.nseg; ++k)
            if (f[k]) {
                stiffnessReady = false;
                throw std::runtime_error("coarsest-level operator is not positive definite (stacked segment " +
                                         std::to_string(k) + ")");
            }
    }
    void prepare_coarsest() {
        if (seg) { seg_prepare(); return; }
        finish_factor();  // buffers below may be reallocated
        LevelData &c = *lv[L];
        hipStream_t s = st();
        const Grid &g = c.g;
        const std::vector<uint8_t> &mask = (L == 0) ? sim->hmask : c.hmask;
        if (L == 0) {
            for (int n = 0; n < g.nnodes; ++n)
                for (int k = 0; k < 3; ++k)
                    if ((mask[n] >> k) & 1)
                        if (sim->hvals[(size_t)k * g.nnodes + n] != 0)
                            throw std::runtime_error("Nonzero Dirichlet constraints currently unsupported");
        }
        std::vector<int> red(3 * (size_t)g.nnodes, -1), redToFull;
        for (int n = 0; n < g.nnodes; ++n) {
            const int iy = (n / g.nn[2]) % g.nn[1];
            for (int k = 0; k < 3; ++k) {
                const bool fixed = ((mask[n] >> k) & 1) || iy >= g.ydet;
                if (!fixed) { red[3 * n + k] = (int)redToFull.size(); redToFull.push_back(3 * n + k); }
            }
        }
        nred = (int)redToFull.size();
        ldA = std::max(64, (nred + 63) / 64 * 64);
        // the solves use the explicit dense L^-1: bound its size (24576^2 doubles = 4.8 GB)
        if (ldA > 24576)
            throw std::runtime_error("direct solve of " + std::to_string(nred) +
                                     " unknowns exceeds the coarse-solver limit (24576); use more coarsening levels");
        // half-bandwidth of the reduced operator in the natural order, in blocks of 64
        int bw = 0;
        for (int n = 0; n < g.nnodes; ++n) {
            const int ix = n / (g.nn[1] * g.nn[2]), iy = (n / g.nn[2]) % g.nn[1], iz = n % g.nn[2];
            for (int k = 0; k < 3; ++k) {
                const int rr = red[3 * (size_t)n + k];
                if (rr < 0) continue;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -1; dz <= 1; ++dz) {
                            const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                            if (jx < 0 || jy < 0 || jz < 0 || jx >= g.nn[0] || jy >= g.nn[1] || jz >= g.nn[2]) continue;
                            const int m = (jx * g.nn[1] + jy) * g.nn[2] + jz;
                            for (int c2 = 0; c2 < 3; ++c2) {
                                const int cc = red[3 * (size_t)m + c2];
                                if (cc >= 0 && cc < rr) bw = std::max(bw, rr - cc);
                            }
                        }
            }
        }
        bandKb = std::max(1, (bw + 63) / 64);
        dRed.resize(red.size());
        dRedToFull.resize(std::max<size_t>(redToFull.size(), 1));
        h2d(dRed.data(), red.data(), red.size(), s);
        h2d(dRedToFull.data(), redToFull.data(), redToFull.size(), s);
        Ainv.resize((size_t)ldA * ldA);
        AinvT.resize((size_t)ldA * ldA);
        band.resize(band_storage(std::max(nred, 1), bandKb));
        invScratch.resize((size_t)(ldA / 64) * 64 * 64);
        ctmp_in.resize(ldA);
        ctmp_out.resize(ldA);
        cholFlag.resize(1);
    }
    // `s`: the stream that queued the coarsest element matrices
    void launch_factor(hipStream_t s) {
        if (seg) { seg_factor(s); return; }
        if (nred == 0) return;  // nothing to solve: the coarse correction is zero
        join_factor();
        if (!fstream) {
            VF_HIP(hipStreamCreateWithFlags(&fstream, hipStreamNonBlocking));
            VF_HIP(hipEventCreateWithFlags(&evGalerkin, hipEventDisableTiming));
            VF_HIP(hipEventCreateWithFlags(&evFactor, hipEventDisableTiming));
        }
        // the coarsest element matrices and the index maps above are complete on the main stream
        VF_HIP(hipEventRecord(evGalerkin, s));
        int device = 0;
        VF_HIP(hipGetDevice(&device));
        factorThread = std::thread([this, device] {
            try {
                VF_HIP(hipSetDevice(device));
                issue_factor();
            } catch (...) {
                factorError = std::current_exception();
            }
        });
        factorPending = true;
    }
    // the factorization's launches on fstream, after the coarsest element matrices (evGalerkin)
    void issue_factor() {
        const Grid &g = lv[L]->g;
        VF_HIP(hipStreamWaitEvent(fstream, evGalerkin, 0));
        auto factor = [&] {
            band_cholesky_inverse(g, lv[L]->Ke.data(), dRed.data(), std::max(nred, 1), bandKb, band.data(),
                                  invScratch.data(), Ainv.data(), AinvT.data(), cholFlag.data(), fstream);
        };
        if (useGraphs) {
            const void *ptrs[7] = {lv[L]->Ke.data(), dRed.data(), band.data(), invScratch.data(),
                                   Ainv.data(), AinvT.data(), cholFlag.data()};
            const int sizes[2] = {nred, bandKb};
            std::string key((const char *)&g, sizeof(Grid));
            key.append((const char *)ptrs, sizeof(ptrs));
            key.append((const char *)sizes, sizeof(sizes));
            if (!factorGraph.exec || key != factorKey) {
                if (factorGraph.exec) VF_HIP(hipGraphExecDestroy(factorGraph.exec));
                if (factorGraph.graph) VF_HIP(hipGraphDestroy(factorGraph.graph));
                factorGraph.exec = nullptr;
                factorGraph.graph = nullptr;
                factorKey.clear();
                hipGraph_t gr = nullptr;
                VF_HIP(hipStreamBeginCapture(fstream, hipStreamCaptureModeThreadLocal));
                try {
                    factor();
                } catch (...) {
                    (void)hipStreamEndCapture(fstream, &gr);
                    if (gr) (void)hipGraphDestroy(gr);
                    throw;
                }
                VF_HIP(hipStreamEndCapture(fstream, &gr));
                factorGraph.graph = gr;
                VF_HIP(hipGraphInstantiate(&factorGraph.exec, gr, nullptr, nullptr, 0));
                factorKey = key;
            }
            VF_HIP(hipGraphLaunch(factorGraph.exec, fstream));
        } else {
            factor();
        }
        VF_HIP(hipEventRecord(evFactor, fstream));
    }

    // ---------------------------------------------------------------- PCG (MultigridSolver.hh:1046-1152)
    int64_t pcg(double *x, const double *b, int64_t maxIter, double tol, int mgIter, int nu, bool fullMG, int prec,
                bool dirichletSatisfied, vf_pcg_callback cb, void *user, vf_pcg_stats *stats) {
        const auto t0 = std::chrono::steady_clock::now();
        sync_masks();
        hipStream_t s = st();
        const Grid &g = lv[0]->g;
        const size_t n0 = 3 * (size_t)g.nnodes;
        pr.resize(n0); ps.resize(n0); pd.resize(n0); pAd.resize(n0);
        double *r = pr.data(), *sv = ps.data(), *d = pd.data(), *Ad = pAd.data();
        double *sc = sim->scalars();
        double *part = sim->red_partials(2 * (size_t)reduction_partials(g) + 2);
        std::vector<double> hx, hr;
        auto callback = [&](int64_t it) {
            if (!cb) return;
            hx.resize(n0); hr.resize(n0);
            d2h(hx.data(), x, n0, s);
            d2h(hr.data(), r, n0, s);
            sync(s);
            if (cb(user, it, hx.data(), hr.data())) throw std::runtime_error("PCG aborted by callback");
        };
        if (L == 0) {
            // direct solve on the fine grid (MultigridSolver.hh:1053-1062); norms over all nodes
            launch_fill<double>(n0, r, 0.0, s);
            apply<double>(0, x, r, b, AM_RESIDUAL, true);
            launch_dot_full(n0, r, r, part, sc + 4, s);
            const double rr = sim->read_scalar(sc + 4);
            launch_dot_full(n0, b, b, part, sc + 7, s);
            const double bb = sim->read_scalar(sc + 7);
            if (stats) { stats->iterations = 0; stats->residual_norm = std::sqrt(rr); stats->rhs_norm = std::sqrt(bb); }
            if (rr < tol * tol * bb) return 0;
            update_stiffness();
            coarse_solve<double>(b, x);
            apply<double>(0, x, r, b, AM_RESIDUAL, true);
            callback(1);
            if (stats) stats->iterations = 1;
            return 1;
        }
        if (!dirichletSatisfied)
            launch_enforce_dirichlet(sim->ndir, sim->dnodes.data(), sim->dnmask.data(), g.nnodes, sim->dvals.data(), x, s);
        launch_dot_active(g, b, b, part, sc + 7, s);
        // r = b - K x; a zero initial guess (every design iteration that does not warm-start) skips the product:
        // r = b on the active rows with the constrained components zeroed, exactly what the product gives
        launch_dot_active(g, x, x, part, sc + 8, s);
        const bool xZero = sim->read_scalar(sc + 8) == 0.0;
        launch_fill<double>(n0, r, 0.0, s);
        if (xZero) {
            launch_copy_active<double, double>(g, 0, b, r, s);
            zero_dirichlet<double>(0, r);
        } else {
            apply<double>(0, x, r, b, AM_RESIDUAL, true);
        }
        launch_dot_active(g, r, r, part, sc + 4, s);
        const double bb = sim->read_scalar(sc + 7);
        double rr = sim->read_scalar(sc + 4);
        if (std::isnan(rr)) throw NanError("NaN encountered");
        int64_t it = 0;
        bool updated = false, first = true;
        const bool flexible = prec == VF_PREC_FP32;
        // Without a host callback the iterate is only needed at the end, so the fused form runs: the
        // preconditioned vector is read in the cycle's precision, x += alpha d is applied inside the next
        // direction update (and once after the loop), and the residual update also writes the FP32
        // right-hand side of the next cycle.
        const bool fused = !cb && mgIter > 0 && nu > 0 && L > 0;
        // the direction updates cover the active rows only (the first one writes them without reading d); the
        // element products also read the detached rows (with zero moduli), which must therefore hold finite values
        if (fused) launch_zero_detached<double>(g, d, s);
        float *r32 = nullptr;
        if (fused && prec == VF_PREC_FP32) {
            ensure_buffers<float>(0);
            r32 = lv[0]->b32.data();
            launch_convert<float, double>(n0, r, r32, s);
        }
        // one PCG iteration, device work only (every launch is on s, no host synchronisation)
        auto body = [&](bool firstIt) {
            if (fused) {
                precondition(r, sv, mgIter, nu, fullMG, prec, true);
                const double *adp = (flexible && !firstIt) ? Ad : nullptr;
                launch_pcg_scalars(sc, 0, s);  // rz_old <- rz
                if (prec == VF_PREC_FP32) {
                    float *s32 = lv[0]->x32.data();
                    zero_dirichlet<float>(0, s32);
                    launch_pcg_dots_s<float>(g, r, s32, adp, part, sc, s);
                    launch_pcg_direction_x<float>(g, s32, d, x, sc, flexible ? 1 : 0, firstIt ? 1 : 0, s);
                } else {
                    zero_dirichlet<double>(0, sv);
                    launch_pcg_dots_s<double>(g, r, sv, adp, part, sc, s);
                    launch_pcg_direction_x<double>(g, sv, d, x, sc, flexible ? 1 : 0, firstIt ? 1 : 0, s);
                }
                if (timeKd) VF_HIP(hipEventRecord(kdEv[0], s));
                apply_dot(d, Ad, part, sc + 3);
                if (timeKd) VF_HIP(hipEventRecord(kdEv[1], s));
                launch_pcg_scalars(sc, 1, s);  // alpha of this iteration
                launch_pcg_residual(g, r, Ad, sc, part, sc + 4, r32, s);
                return;
            }
            if (mgIter > 0) precondition(r, sv, mgIter, nu, fullMG, prec);
            else d2d(sv, r, n0, s);
            zero_dirichlet<double>(0, sv);
            launch_pcg_scalars(sc, 0, s);  // rz_old <- rz
            if (flexible && !firstIt) launch_dot2_active(g, r, sv, Ad, part, sc, s);  // sc[0]=r.s, sc[1]=Ad_prev.s
            else launch_dot_active(g, r, sv, part, sc, s);
            if (firstIt) d2d(d, sv, n0, s);
            else launch_pcg_direction(g, sv, d, sc, flexible ? 1 : 0, s);
            apply<double>(0, d, Ad, nullptr, AM_SET, true);
            launch_dot_active(g, d, Ad, part, sc + 3, s);
            launch_pcg_scalars(sc, 1, s);  // alpha of this iteration
            launch_pcg_update(g, x, r, d, Ad, sc, part, sc + 4, s);
        };
        // With useGraphs, the steady-state iteration (second and later) is captured once into a HIP graph and
        // replayed: the first iteration runs eagerly (it also builds the hierarchy, tunes kernels and
        // allocates work buffers, none of which may happen under capture).
        // The capture is host work only: it is done right after the eager first iteration has been queued,
        // while the device runs it, instead of at the second iteration with the device idle.
        GraphGuard graph;
        auto capture = [&] {
            retire_graph();
            hipGraph_t gr = nullptr;
            VF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                body(false);
            } catch (...) {
                (void)hipStreamEndCapture(s, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            VF_HIP(hipStreamEndCapture(s, &gr));
            graph.graph = gr;
            VF_HIP(hipGraphInstantiate(&graph.exec, gr, nullptr, nullptr, 0));
        };
        while (it < maxIter && rr > tol * tol * bb) {
            ++it;
            // the reference rebuilds the hierarchy at the first iteration of every solve
            // (MultigridSolver.hh:1104-1107); here it is rebuilt only if the moduli changed since the last
            // build (every change of densities, law, material or mask layers bumps the simulator's version)
            if (mgIter > 0 && !updated) { ensure_stiffness(); updated = true; }
            if (useGraphs && !first && s != nullptr) {
                if (!graph.exec) capture();
                VF_HIP(hipGraphLaunch(graph.exec, s));
            } else {
                body(first);
                if (useGraphs && first && s != nullptr && it < maxIter) capture();
            }
            rr = sim->read_scalar(sc + 4);
            if (timeKd && fused) {  // the stream is idle here: both events of this iteration have completed
                float ms = 0;
                VF_HIP(hipEventElapsedTime(&ms, kdEv[0], kdEv[1]));
                kdMs += ms;
                ++kdCount;
            }
            if (std::isnan(rr)) {
                double v[6];
                d2h(v, sc, 6, s);
                sync(s);
                throw NanError("NaN encountered at iteration" + std::to_string(it) + " (<r,s> " + std::to_string(v[0]) +
                               ", <d,Ad> " + std::to_string(v[3]) + ", <r,r> " + std::to_string(v[4]) + ")");
            }
            callback(it);
            first = false;
        }
        if (fused && it > 0) launch_pcg_finish_x(g, x, d, sc, s);
        if (graph.exec) {
            retire_graph();
            retiredGraph.exec = graph.exec;
            retiredGraph.graph = graph.graph;
            graph.exec = nullptr;
            graph.graph = nullptr;
        }
        if (stats) {
            stats->iterations = it;
            stats->residual_norm = std::sqrt(rr);
            stats->rhs_norm = std::sqrt(bb);
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return it;
    }
};

template <> double *vf_mg::X<double>(int l) { return l == 0 && x0v ? x0v : lv[l]->x.data(); }
template <> double *vf_mg::B<double>(int l) { return l == 0 && b0v ? b0v : lv[l]->b.data(); }
template <> double *vf_mg::R<double>(int l) {
    if (lv[l]->r.empty()) lv[l]->r.resize(3 * (size_t)lv[l]->g.nnodes);
    return lv[l]->r.data();
}
template <> float *vf_mg::X<float>(int l) { return lv[l]->x32.data(); }
template <> float *vf_mg::B<float>(int l) { return lv[l]->b32.data(); }
template <> float *vf_mg::R<float>(int l) { return lv[l]->r32.data(); }
template <> const double *vf_mg::Wp<double>() const { return sim->Wtab.data(); }
template <> const float *vf_mg::Wp<float>() const { return sim->Wtab32.data(); }
template <> const double *vf_mg::Fp<double>() const { return sim->Ftab.data(); }
template <> const float *vf_mg::Fp<float>() const { return sim->Ftab32.data(); }
template <> const double *vf_mg::K0p<double>() const { return sim->K0.data(); }
template <> const float *vf_mg::K0p<float>() const { return sim->K032.data(); }
template <> const double *vf_mg::Ep<double>() const { return sim->E.data(); }
template <> const float *vf_mg::Ep<float>() const { return sim->E32.data(); }
template <> const double *vf_mg::Chatp<double>() const { return dChat.data(); }
template <> const float *vf_mg::Chatp<float>() const { return dChat32.data(); }
template <> const double *vf_mg::Cwp<double>() const { return dCw.data(); }
template <> const float *vf_mg::Cwp<float>() const { return dCw32.data(); }
template <> const double *vf_mg::blocksp<double>(int l) const {
    if (!blocksReady[0]) throw std::runtime_error("internal: FP64 block rows used before assembly");
    return lv[l]->blocks.data();
}
template <> const float *vf_mg::blocksp<float>(int l) const {
    if (!blocksReady[1]) throw std::runtime_error("internal: FP32 block rows used before assembly");
    return lv[l]->blocks32.data();
}
template <> DBuf<double> &vf_mg::block_buf<double>(int l) { return lv[l]->blocks; }
template <> DBuf<float> &vf_mg::block_buf<float>(int l) { return lv[l]->blocks32; }
void vf_mg::apply_dot(const double *d, double *Ad, double *partials, double *dAd) {
    if (ku_family<double>() == 2) {
        launch_fine_apply_walsh<double>(lv[0]->g, Wp<double>(), Ep<double>(), d, Ad, nullptr, dm(0), AM_SET, st(),
                                        partials, dAd);
    } else {
        apply<double>(0, d, Ad, nullptr, AM_SET, true);
        launch_dot_active(lv[0]->g, d, Ad, partials, dAd, st());
    }
}

// =================================================================================================
// C ABI
// =================================================================================================
#define VF_API_BEGIN try {
#define VF_API_END                                                                                       \
    return VF_OK;                                                                                        \
    }                                                                                                    \
    catch (const NanError &e) { g_err = e.what(); return VF_ERR_NAN; }                                    \
    catch (const HipError &e) { g_err = e.what(); return VF_ERR_HIP; }                                    \
    catch (const std::invalid_argument &e) { g_err = e.what(); return VF_ERR_ARG; }                       \
    catch (const std::exception &e) { g_err = e.what(); return VF_ERR_RUNTIME; }                          \
    catch (...) { g_err = "unknown error"; return VF_ERR_RUNTIME; }

namespace {
void need(bool ok, const char *msg) {
    if (!ok) throw std::invalid_argument(msg);
}

// Run `fn(dev_in..., dev_out)` with staging for host pointers.
// host-pointer staging of the planar entry points (slots 0..2 of the simulator's staging buffers)
struct P2Stage {
    vf::planar::PSim &s;
    int where;
    const double *in(const double *p, size_t n, int slot) {
        if (where == VF_DEVICE) return p;
        double *d = s.stage(slot, n);
        h2d(d, p, n, s.stream);
        return d;
    }
    double *inout(double *p, size_t n, int slot) { return const_cast<double *>(in(p, n, slot)); }
    double *out(double *p, size_t n, int slot) { return where == VF_DEVICE ? p : s.stage(slot, n); }
    void finish(double *p, const double *d, size_t n) {
        if (where == VF_HOST) d2h(p, d, n, s.stream);
        sync(s.stream);
    }
};

struct Staged {
    vf_sim *sim;
    int where;
    const double *in(const double *p, size_t n, double *stage) {
        if (where == VF_DEVICE) return p;
        h2d(stage, p, n, sim->stream);
        return stage;
    }
    double *out(double *p, double *stage) { return where == VF_DEVICE ? p : stage; }
    void finish(double *p, const double *stage, size_t n) {
        if (where == VF_HOST) { d2h(p, stage, n, sim->stream); sync(sim->stream); }
    }
};
}  // namespace

// host-pointer staging for the stand-alone filter/optimiser entry points
static hipStream_t as_stream(void *s) { return (hipStream_t)s; }

template <typename F>
static void with_buffers(size_t nin, size_t nout, const double *in, double *out, int where, hipStream_t s, F &&fn,
                         const double *in2 = nullptr) {
    if (where == VF_DEVICE) { fn(in, in2, out); return; }
    DBuf<double> a(nin), b(nout), c(in2 ? nin : 0);
    h2d(a.data(), in, nin, s);
    if (in2) h2d(c.data(), in2, nin, s);
    fn(a.data(), in2 ? c.data() : nullptr, b.data());
    d2h(out, b.data(), nout, s);
    sync(s);
}

// ---- allocation-free mean: its scratch lives in a file-scope device array -----------------------
// vf_mean_noalloc keeps its partial sums and its sum in `scratch` (one copy per device, held until the
// process ends) instead of allocating them, so it can run where hipMalloc cannot (graph capture, serving
// loops). Layout: NOALLOC_PARTIALS partial sums (one per TPB*ITEMS = 2048 inputs), then the sum at
// NOALLOC_PARTIALS. The array's name is a host-side stand-in; host code uses the address that
// hipGetSymbolAddress returns for the current device.
namespace {
constexpr int NOALLOC_PARTIALS = 32768;
constexpr int NOALLOC_SLOTS = NOALLOC_PARTIALS + 32;
constexpr int64_t NOALLOC_MAX_N = (int64_t)NOALLOC_PARTIALS * 2048;
}  // namespace
static __device__ __align__(256) double scratch[NOALLOC_SLOTS];
namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // device address of `scratch`, per device id

double *scratch_on_current_device() {
    int dev = 0;
    VF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        VF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        need(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

const char *vf_last_error(void) { return g_err.c_str(); }

int vf_malloc(void **ptr, size_t bytes) {
    VF_API_BEGIN
    need(ptr != nullptr, "null pointer");
    *ptr = nullptr;
    if (bytes) VF_HIP(hipMalloc(ptr, bytes));
    VF_API_END
}

int vf_free(void *ptr) {
    VF_API_BEGIN
    if (ptr) VF_HIP(hipFree(ptr));
    VF_API_END
}

int vf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    VF_API_BEGIN
    need(kind >= 0 && kind <= 2, "unknown copy kind");
    if (!bytes) return VF_OK;
    need(dst && src, "null pointer");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    VF_HIP(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
    if (kind != 2) sync((hipStream_t)stream);
    VF_API_END
}

int vf_stream_create(void **stream) {
    VF_API_BEGIN
    need(stream != nullptr, "null argument");
    hipStream_t s = nullptr;
    VF_HIP(hipStreamCreateWithFlags(&s, hipStreamDefault));
    *stream = (void *)s;
    VF_API_END
}

int vf_stream_destroy(void *stream) {
    VF_API_BEGIN
    if (stream) VF_HIP(hipStreamDestroy((hipStream_t)stream));
    VF_API_END
}

int vf_stream_sync(void *stream) {
    VF_API_BEGIN
    sync((hipStream_t)stream);
    VF_API_END
}
const char *vf_version(void) { return "voxelfem-mi355x 0.1 (gfx950)"; }

int vf_set_kernel_variant(int variant) {
    VF_API_BEGIN
    need(variant >= 0 && variant <= 5,
         "kernel variant must be 0 (autotuned), 1 (register-gather), 2 (LDS-tiled), 3 (Walsh K u, LDS-tiled smoother), "
         "4 (Walsh K u, structured register-gather smoother) or 5 (as 3, level 1 on the fly on the matrix cores)");
    g_kernel_variant = variant;
    VF_API_END
}

int vf_device_count(int *count) {
    VF_API_BEGIN
    need(count != nullptr, "null count");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    VF_API_END
}

int vf_sim_create(vf_sim **out, const int64_t ne[3], const double minCorner[3], const double maxCorner[3],
                  void *stream) {
    VF_API_BEGIN
    need(out && ne && minCorner && maxCorner, "null argument");
    *out = nullptr;
    auto s = std::make_unique<vf_sim>();
    s->stream = (hipStream_t)stream;
    for (int d = 0; d < 3; ++d) {
        need(ne[d] >= 1, "grid must have at least one element per dimension");
        s->ne[d] = ne[d];
        s->lo[d] = minCorner[d];
        s->hi[d] = maxCorner[d];
        s->h[d] = (maxCorner[d] - minCorner[d]) / (double)ne[d];
        need(s->h[d] > 0, "domain must have positive extent");
    }
    need((ne[0] + 1) * (ne[1] + 1) * (ne[2] + 1) * 3 < INT_MAX, "grid too large for 32-bit node indexing");
    s->g = make_grid(ne);
    s->K0.resize(576);
    s->K032.resize(576);
    s->Ftab.resize(128);
    s->Ftab32.resize(128);
    s->Wtab.resize(128);
    s->Wtab32.resize(128);
    s->rho.resize(s->nel());
    s->E.resize(s->nel());
    s->E32.resize(s->nel());
    launch_fill<double>(s->nel(), s->rho.data(), 0.0, s->stream);
    s->hmask.assign(s->nn(), 0);
    s->hvals.assign(3 * s->nn(), 0.0);
    upload_dirichlet(s->hmask, s->dmask, s->dnodes, s->dnmask, s->ndir, s->stream);
    s->dvals.resize(3);
    s->hforce.assign(3 * s->nn(), 0.0);
    s->update_K0();
    s->update_E();
    sync(s->stream);
    *out = s.release();
    VF_API_END
}

int vf_sim_create_2d(vf_sim **out, const int64_t ne[2], const double minCorner[2], const double maxCorner[2],
                     void *stream) {
    VF_API_BEGIN
    need(out && ne && minCorner && maxCorner, "null argument");
    *out = nullptr;
    auto s = std::make_unique<vf_sim>();
    s->stream = (hipStream_t)stream;
    s->p2 = std::make_unique<vf::planar::PSim>(ne, minCorner, maxCorner, s->stream);
    for (int d = 0; d < 2; ++d) {
        s->ne[d] = ne[d];
        s->lo[d] = minCorner[d];
        s->hi[d] = maxCorner[d];
        s->h[d] = s->p2->h[d];
    }
    s->ne[2] = 0;
    s->lo[2] = s->hi[2] = s->h[2] = 0;
    *out = s.release();
    VF_API_END
}

int vf_sim_dimension(vf_sim *sim, int *dim) {
    VF_API_BEGIN
    need(sim && dim, "null argument");
    *dim = sim->p2 ? 2 : 3;
    VF_API_END
}

int vf_sim_destroy(vf_sim *sim) {
    VF_API_BEGIN
    if (sim) { (void)hipStreamSynchronize(sim->stream); delete sim; }
    VF_API_END
}

int vf_sim_set_stream(vf_sim *sim, void *stream) {
    VF_API_BEGIN
    need(sim, "null sim");
    sync(sim->stream);
    sim->stream = (hipStream_t)stream;
    if (sim->p2) sim->p2->stream = sim->stream;
    VF_API_END
}

int vf_sim_set_material(vf_sim *sim, double young, double poisson) {
    VF_API_BEGIN
    need(sim, "null sim");
    need(young > 0 && poisson > -1 && poisson < 0.5, "invalid isotropic material");
    if (sim->p2) { sim->p2->set_material(young, poisson); return VF_OK; }
    sim->young = young;
    sim->poisson = poisson;
    sim->update_K0();
    ++sim->version;
    ++sim->baseVersion;
    VF_API_END
}

int vf_sim_set_law(vf_sim *sim, int law, double E0, double Emin, double gamma, double q) {
    VF_API_BEGIN
    need(sim, "null sim");
    need(law == VF_SIMP || law == VF_RAMP, "unknown interpolation law");
    if (sim->p2) { sim->p2->set_law(law, E0, Emin, gamma, q); return VF_OK; }
    sim->law = law;
    sim->E0 = E0;
    sim->Emin = Emin;
    sim->gamma = gamma;
    sim->q = q;
    sim->update_E();
    VF_API_END
}

int vf_sim_set_densities(vf_sim *sim, const double *rho, int where) {
    VF_API_BEGIN
    need(sim && rho, "null argument");
    if (sim->p2) {
        sim->p2->set_densities(rho, where == VF_DEVICE);
        if (where == VF_HOST) sync(sim->stream);
        return VF_OK;
    }
    if (where == VF_DEVICE) d2d(sim->rho.data(), rho, sim->nel(), sim->stream);
    else h2d(sim->rho.data(), rho, sim->nel(), sim->stream);
    sim->update_E();
    if (where == VF_HOST) sync(sim->stream);
    VF_API_END
}

int vf_sim_get_densities(vf_sim *sim, double *rho, int where) {
    VF_API_BEGIN
    need(sim && rho, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        if (where == VF_DEVICE) d2d(rho, p2.rho.data(), p2.nel(), p2.stream);
        else { d2h(rho, p2.rho.data(), p2.nel(), p2.stream); sync(p2.stream); }
        return VF_OK;
    }
    if (where == VF_DEVICE) d2d(rho, sim->rho.data(), sim->nel(), sim->stream);
    else { d2h(rho, sim->rho.data(), sim->nel(), sim->stream); sync(sim->stream); }
    VF_API_END
}

int vf_sim_get_young(vf_sim *sim, double *E, int where) {
    VF_API_BEGIN
    need(sim && E, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        if (where == VF_DEVICE) d2d(E, p2.E.data(), p2.nel(), p2.stream);
        else { d2h(E, p2.E.data(), p2.nel(), p2.stream); sync(p2.stream); }
        return VF_OK;
    }
    if (where == VF_DEVICE) d2d(E, sim->E.data(), sim->nel(), sim->stream);
    else { d2h(E, sim->E.data(), sim->nel(), sim->stream); sync(sim->stream); }
    VF_API_END
}

int vf_sim_get_K0(vf_sim *sim, double *K0) {
    VF_API_BEGIN
    need(sim && K0, "null argument");
    if (sim->p2) { std::copy(sim->p2->K0h, sim->p2->K0h + 64, K0); return VF_OK; }
    std::copy(sim->K0h, sim->K0h + 576, K0);
    VF_API_END
}

int vf_sim_set_dirichlet(vf_sim *sim, const uint8_t *mask, const double *values) {
    VF_API_BEGIN
    need(sim && mask, "null argument");
    if (sim->p2) { sim->p2->set_dirichlet(mask, values); return VF_OK; }
    const size_t nn = sim->nn();
    for (size_t n = 0; n < nn; ++n) sim->hmask[n] = mask[n] & 7;
    for (size_t i = 0; i < 3 * nn; ++i) sim->hvals[i] = values ? values[i] : 0.0;
    ++sim->bcVersion;
    upload_dirichlet(sim->hmask, sim->dmask, sim->dnodes, sim->dnmask, sim->ndir, sim->stream);
    std::vector<double> v(3 * std::max(sim->ndir, 1), 0.0);
    int k = 0;
    for (size_t n = 0; n < nn; ++n) {
        if (!sim->hmask[n]) continue;
        for (int c = 0; c < 3; ++c) v[3 * k + c] = sim->hvals[c * nn + n];
        ++k;
    }
    sim->dvals.resize(v.size());
    h2d(sim->dvals.data(), v.data(), v.size(), sim->stream);
    sync(sim->stream);
    VF_API_END
}

int vf_sim_set_forces(vf_sim *sim, const double *f) {
    VF_API_BEGIN
    need(sim && f, "null argument");
    if (sim->p2) { sim->p2->set_forces(f); return VF_OK; }
    std::copy(f, f + 3 * sim->nn(), sim->hforce.begin());
    sim->hasForce = std::any_of(sim->hforce.begin(), sim->hforce.end(), [](double v) { return v != 0.0; });
    sim->force.resize(3 * sim->nn());
    h2d(sim->force.data(), f, 3 * sim->nn(), sim->stream);
    sync(sim->stream);
    VF_API_END
}

int vf_sim_set_gravity(vf_sim *sim, const double g[3]) {
    VF_API_BEGIN
    need(sim && g, "null argument");
    if (sim->p2) { sim->p2->grav[0] = g[0]; sim->p2->grav[1] = g[1]; return VF_OK; }
    for (int c = 0; c < 3; ++c) sim->grav[c] = g[c];
    VF_API_END
}

int vf_sim_set_mask_layer(vf_sim *sim, int64_t layer) {
    VF_API_BEGIN
    need(sim, "null sim");
    if (sim->p2) { sim->p2->set_mask_height(sim->p2->h[1] * (double)layer, true); return VF_OK; }
    sim->set_mask_height(sim->h[1] * (double)layer, true);
    VF_API_END
}

int vf_sim_get_mask_info(vf_sim *sim, int64_t *firstMasked, int64_t *firstDetached, double *height) {
    VF_API_BEGIN
    need(sim, "null sim");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        if (firstMasked) *firstMasked = p2.firstMasked;
        if (firstDetached) *firstDetached = p2.firstDetached;
        if (height) *height = p2.maskHeight;
        return VF_OK;
    }
    if (firstMasked) *firstMasked = sim->firstMasked;
    if (firstDetached) *firstDetached = sim->firstDetached;
    if (height) *height = sim->maskHeight;
    VF_API_END
}

int vf_sim_apply_K(vf_sim *sim, const double *u, double *y, int where) {
    VF_API_BEGIN
    need(sim && u && y, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        const size_t n = p2.ndof();
        P2Stage st{p2, where};
        const double *du = st.in(u, n, 0);
        double *dy = st.out(y, n, 1);
        p2.apply_K(du, dy);
        st.finish(y, dy, n);
        return VF_OK;
    }
    sim->refresh_grid();
    const size_t n = 3 * sim->nn();
    Staged st{sim, where};
    const double *du = st.in(u, n, sim->stage_in(n));
    double *dy = st.out(y, sim->stage_out(n));
    sim_apply_K<double>(sim, du, dy, sim->stream);
    st.finish(y, dy, n);
    VF_API_END
}

int vf_sim_build_load(vf_sim *sim, double *f, int where) {
    VF_API_BEGIN
    need(sim && f, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        const size_t n = p2.ndof();
        P2Stage st{p2, where};
        double *df = st.out(f, n, 1);
        p2.build_load(df);
        st.finish(f, df, n);
        return VF_OK;
    }
    sim->refresh_grid();
    const size_t n = 3 * sim->nn();
    Staged st{sim, where};
    double *df = st.out(f, sim->stage_out(n));
    if (sim->hasForce) d2d(df, sim->force.data(), n, sim->stream);
    else launch_fill<double>(n, df, 0.0, sim->stream);
    if (sim->grav[0] != 0 || sim->grav[1] != 0 || sim->grav[2] != 0)
        launch_self_weight(sim->g, sim->rho.data(), sim->grav, sim->h[0] * sim->h[1] * sim->h[2], df, sim->stream);
    st.finish(f, df, n);
    VF_API_END
}

static int grad_impl(vf_sim *sim, const double *u, double *g, int where, int accumulate) {
    VF_API_BEGIN
    need(sim && u && g, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        const size_t n = p2.ndof(), ne = p2.nel();
        P2Stage st{p2, where};
        const double *du = st.in(u, n, 0);
        double *dg = where == VF_DEVICE ? g : p2.stage(1, ne);
        if (where == VF_HOST && accumulate) h2d(dg, g, ne, p2.stream);
        p2.compliance_gradient(du, dg, accumulate != 0);
        st.finish(g, dg, ne);
        return VF_OK;
    }
    sim->refresh_grid();
    const size_t n = 3 * sim->nn(), ne = sim->nel();
    Staged st{sim, where};
    const double *du = st.in(u, n, sim->stage_in(n));
    double *dg = where == VF_DEVICE ? g : sim->stage_out(ne);
    if (where == VF_HOST && accumulate) h2d(dg, g, ne, sim->stream);
    launch_compliance_gradient(sim->g, sim->K0.data(), sim->rho.data(), du, dg, sim->law, sim->E0, sim->Emin,
                               sim->gamma, sim->q, sim->grav, sim->h[0] * sim->h[1] * sim->h[2], accumulate,
                               sim->stream);
    st.finish(g, dg, ne);
    VF_API_END
}

int vf_sim_compliance_gradient(vf_sim *sim, const double *u, double *g, int where) { return grad_impl(sim, u, g, where, 0); }
int vf_sim_accumulate_compliance_gradient(vf_sim *sim, const double *u, double *g, int where) {
    return grad_impl(sim, u, g, where, 1);
}

int vf_sim_energy_density(vf_sim *sim, const double *u, double *out, int where) {
    VF_API_BEGIN
    need(sim && u && out, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        const size_t n = p2.ndof(), ne = p2.nel();
        P2Stage st{p2, where};
        const double *du = st.in(u, n, 0);
        double *dout = st.out(out, ne, 1);
        p2.energy_density(du, dout);
        st.finish(out, dout, ne);
        return VF_OK;
    }
    const size_t n = 3 * sim->nn(), ne = sim->nel();
    Staged st{sim, where};
    const double *du = st.in(u, n, sim->stage_in(n));
    double *dout = st.out(out, sim->stage_out(ne));
    launch_energy_density(sim->g, sim->K0.data(), sim->E.data(), du, dout, sim->stream);
    st.finish(out, dout, ne);
    VF_API_END
}

int vf_sim_dot(vf_sim *sim, const double *a, const double *b, int masked, int where, double *out) {
    VF_API_BEGIN
    need(sim && a && b && out, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        const size_t n = p2.ndof();
        P2Stage st{p2, where};
        const double *da = st.in(a, n, 0);
        const double *db = st.in(b, n, 1);
        *out = p2.dot(da, db, masked != 0);
        return VF_OK;
    }
    sim->refresh_grid();
    const size_t n = 3 * sim->nn();
    Staged st{sim, where};
    const double *da = st.in(a, n, sim->stage_in(n));
    const double *db = st.in(b, n, sim->stage_aux(n));
    double *sc = sim->scalars();
    if (masked) {
        launch_dot_active(sim->g, da, db, sim->red_partials(reduction_partials(sim->g)), sc + 9, sim->stream);
        *out = sim->read_scalar(sc + 9);
    } else {
        launch_dot_full(n, da, db, sim->red_partials(reduction_partials_full(n)), sc + 9, sim->stream);
        *out = sim->read_scalar(sc + 9);
    }
    VF_API_END
}

int vf_sim_constant_strain_load(vf_sim *sim, const double eps[9], double *F, int where) {
    VF_API_BEGIN
    need(sim && eps && F, "null argument");
    if (sim->p2) {
        auto &p2 = *sim->p2;
        need(std::fabs(eps[1] - eps[2]) <= 1e-12 * (std::fabs(eps[1]) + std::fabs(eps[2]) + 1e-300),
             "constantStrainLoad: the strain must be a symmetric 2x2 matrix");
        const size_t n = p2.ndof();
        P2Stage st{p2, where};
        double *dF = st.out(F, n, 1);
        p2.constant_strain_load(eps, dF);
        st.finish(F, dF, n);
        return VF_OK;
    }
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            need(std::fabs(eps[3 * a + b] - eps[3 * b + a]) <= 1e-12 * (std::fabs(eps[3 * a + b]) + std::fabs(eps[3 * b + a]) + 1e-300),
                 "constantStrainLoad: the strain must be a symmetric 3x3 matrix");
    // isotropic elasticity tensor of the material (setETensor): sigma = lambda tr(eps) I + 2 mu eps
    const double E = sim->young, nu = sim->poisson;
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    const double tr = eps[0] + eps[4] + eps[8];
    const double V = sim->h[0] * sim->h[1] * sim->h[2];
    StressCoef sc;
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
            const double sigma = (c == d ? lam * tr : 0.0) + 2 * mu * eps[3 * c + d];
            sc.w[c][d] = sigma * V / (4.0 * sim->h[d]);
        }
    const size_t n = 3 * sim->nn();
    Staged st{sim, where};
    double *dF = st.out(F, sim->stage_out(n));
    launch_constant_strain_load(sim->g, sim->rho.data(), sc, dF, sim->stream);
    st.finish(F, dF, n);
    VF_API_END
}

int vf_sim_sample_nodal_field(vf_sim *sim, const double *u, int64_t ncomp, const double *p, int64_t np, double *out,
                              int where) {
    VF_API_BEGIN
    need(sim && u && p && out, "null argument");
    need(ncomp >= 1 && np >= 0 && np < INT_MAX / 3, "invalid sizes");
    if (np == 0) return VF_OK;
    if (sim->p2) {
        auto &p2 = *sim->p2;
        hipStream_t s = p2.stream;
        const size_t nu = (size_t)ncomp * p2.nn(), npo = (size_t)ncomp * np;
        DBuf<double> hu, hp, ho;
        const double *du = u, *dp = p;
        double *dout = out;
        if (where == VF_HOST) {
            hu.resize(nu); hp.resize(2 * (size_t)np); ho.resize(npo);
            h2d(hu.data(), u, nu, s);
            h2d(hp.data(), p, 2 * (size_t)np, s);
            du = hu.data(); dp = hp.data(); dout = ho.data();
        }
        DBuf<int> bad(1);
        launch_fill<int>(1, bad.data(), 0, s);
        p2.sample(du, (int)ncomp, dp, (int)np, dout, bad.data());
        int flag = 0;
        d2h(&flag, bad.data(), 1, s);
        if (where == VF_HOST) d2h(out, dout, npo, s);
        sync(s);
        if (flag) throw std::runtime_error("Point out of bounds");
        return VF_OK;
    }
    hipStream_t s = sim->stream;
    const size_t nu = (size_t)ncomp * sim->nn(), npo = (size_t)ncomp * np;
    DBuf<double> hu, hp, ho;
    const double *du = u, *dp = p;
    double *dout = out;
    if (where == VF_HOST) {
        hu.resize(nu); hp.resize(3 * (size_t)np); ho.resize(npo);
        h2d(hu.data(), u, nu, s);
        h2d(hp.data(), p, 3 * (size_t)np, s);
        du = hu.data(); dp = hp.data(); dout = ho.data();
    }
    DBuf<int> bad(1);
    launch_fill<int>(1, bad.data(), 0, s);
    launch_sample_nodal(sim->g, sim->lo, sim->hi, sim->h, du, (int)ncomp, dp, (int)np, dout, bad.data(), s);
    int flag = 0;
    d2h(&flag, bad.data(), 1, s);
    if (where == VF_HOST) d2h(out, dout, npo, s);
    sync(s);
    if (flag) throw std::runtime_error("Point out of bounds");
    VF_API_END
}

int vf_sim_set_stretchings(vf_sim *sim, const double h[3]) {
    VF_API_BEGIN
    need(sim && h, "null argument");
    if (sim->p2) { sim->p2->set_stretchings(h); return VF_OK; }
    for (int d = 0; d < 3; ++d) need(h[d] > 0, "stretchings must be positive");
    std::copy(h, h + 3, sim->h);
    sim->update_K0();
    ++sim->version;
    ++sim->baseVersion;
    VF_API_END
}

int vf_sim_get_stretchings(vf_sim *sim, double h[3]) {
    VF_API_BEGIN
    need(sim && h, "null argument");
    if (sim->p2) { h[0] = sim->p2->h[0]; h[1] = sim->p2->h[1]; h[2] = 0.0; return VF_OK; }
    std::copy(sim->h, sim->h + 3, h);
    VF_API_END
}

// ---------------------------------------------------------------------------------------------- MG
int vf_mg_create(vf_mg **out, vf_sim *fine, int64_t levels) {
    VF_API_BEGIN
    need(out && fine, "null argument");
    *out = nullptr;
    if (fine->p2) {
        *out = new vf_mg(fine, std::make_unique<vf::planar::PMG>(fine->p2.get(), (int)levels));
        return VF_OK;
    }
    *out = new vf_mg(fine, (int)levels);
    VF_API_END
}

int vf_mg_destroy(vf_mg *mg) {
    VF_API_BEGIN
    if (mg) { (void)hipStreamSynchronize(mg->st()); delete mg; }
    VF_API_END
}

int vf_mg_num_levels(vf_mg *mg, int64_t *levels) {
    VF_API_BEGIN
    need(mg && levels, "null argument");
    *levels = mg->L;
    VF_API_END
}

int vf_mg_level_nodes(vf_mg *mg, int64_t l, int64_t nn[3]) {
    VF_API_BEGIN
    need(mg && nn, "null argument");
    need(l >= 0 && l <= mg->L, "level out of range");
    if (mg->p2) {
        nn[0] = mg->p2->lv[l]->g.nn[0];
        nn[1] = mg->p2->lv[l]->g.nn[1];
        nn[2] = 1;
        return VF_OK;
    }
    for (int d = 0; d < 3; ++d) nn[d] = mg->lv[l]->g.nn[d];
    VF_API_END
}

int vf_mg_level_dirichlet(vf_mg *mg, int64_t l, uint8_t *mask) {
    VF_API_BEGIN
    need(mg && mask, "null argument");
    need(l >= 0 && l <= mg->L, "level out of range");
    if (mg->p2) {
        const auto &m = l == 0 ? mg->p2->sim->hmask : mg->p2->lv[l]->hmask;
        std::copy(m.begin(), m.end(), mask);
        return VF_OK;
    }
    const auto &m = l == 0 ? mg->sim->hmask : mg->lv[l]->hmask;
    std::copy(m.begin(), m.end(), mask);
    VF_API_END
}

int vf_mg_set_symmetric_gs(vf_mg *mg, int symmetric) {
    VF_API_BEGIN
    need(mg, "null mg");
    mg->symmetricGS = symmetric != 0;
    if (mg->p2) mg->p2->symmetricGS = symmetric != 0;
    VF_API_END
}

int vf_mg_update_stiffness(vf_mg *mg) {
    VF_API_BEGIN
    need(mg, "null mg");
    if (mg->p2) { mg->p2->update_stiffness(); return VF_OK; }
    // asynchronous on the solver's stream: later calls are ordered after it, and a failed factorisation
    // is reported by the first coarse solve
    mg->update_stiffness();
    VF_API_END
}

int vf_mg_finish_update(vf_mg *mg) {
    VF_API_BEGIN
    need(mg, "null mg");
    if (mg->p2) { sync(mg->p2->st()); return VF_OK; }
    mg->finish_factor();
    VF_API_END
}

int vf_mg_set_mask_layer(vf_mg *mg, int64_t layer) {
    VF_API_BEGIN
    need(mg, "null mg");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.sim->set_mask_height(m2.sim->h[1] * (double)layer, true);
        m2.sync_masks();
        m2.stiffReady = false;
        return VF_OK;
    }
    mg->sim->set_mask_height(mg->sim->h[1] * (double)layer, true);
    mg->sync_masks();
    mg->stiffnessReady = false;
    VF_API_END
}

static size_t level_n(vf_mg *mg, int64_t l) { return 3 * (size_t)mg->lv[l]->g.nnodes; }

int vf_mg_apply_K(vf_mg *mg, int64_t l, const double *u, double *y, int where) {
    VF_API_BEGIN
    need(mg && u && y, "null argument");
    need(l >= 0 && l <= mg->L, "level out of range");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.sync_masks();
        if (l >= 1) m2.ensure_stiffness();
        const size_t n = m2.level_n((int)l);
        P2Stage st{*m2.sim, where};
        const double *du = st.in(u, n, 0);
        double *dy = st.out(y, n, 1);
        m2.apply_K((int)l, du, dy, true, false);
        st.finish(y, dy, n);
        return VF_OK;
    }
    mg->sync_masks();
    if (l >= 1) mg->ensure_blocks<double>();
    const size_t n = level_n(mg, l);
    Staged st{mg->sim, where};
    const double *du = st.in(u, n, mg->sim->stage_in(n));
    double *dy = st.out(y, mg->sim->stage_out(n));
    launch_fill<double>(n, dy, 0.0, mg->st());
    mg->apply<double>((int)l, du, dy, nullptr, AM_SET, false);
    st.finish(y, dy, n);
    VF_API_END
}

int vf_mg_residual(vf_mg *mg, int64_t l, const double *u, const double *b, double *r, int where) {
    VF_API_BEGIN
    need(mg && u && b && r, "null argument");
    need(l >= 0 && l <= mg->L, "level out of range");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.sync_masks();
        if (l >= 1) m2.ensure_stiffness();
        const size_t n = m2.level_n((int)l);
        P2Stage st{*m2.sim, where};
        const double *du = st.in(u, n, 0);
        const double *db = st.in(b, n, 1);
        double *dr = st.out(r, n, 2);
        launch_fill<double>(n, dr, 0.0, m2.st());
        m2.residual((int)l, du, db, dr);
        st.finish(r, dr, n);
        return VF_OK;
    }
    mg->sync_masks();
    if (l >= 1) mg->ensure_blocks<double>();
    const size_t n = level_n(mg, l);
    Staged st{mg->sim, where};
    const double *du = st.in(u, n, mg->sim->stage_in(n));
    const double *db = st.in(b, n, mg->sim->stage_aux(n));
    double *dr = st.out(r, mg->sim->stage_out(n));
    launch_fill<double>(n, dr, 0.0, mg->st());
    mg->apply<double>((int)l, du, dr, db, AM_RESIDUAL, true);
    st.finish(r, dr, n);
    VF_API_END
}

int vf_mg_smooth(vf_mg *mg, int64_t l, double *u, const double *b, int forward, int where) {
    VF_API_BEGIN
    need(mg && u && b, "null argument");
    need(l >= 0 && l < mg->L, "smoothing needs a non-coarsest level");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.sync_masks();
        if (l >= 1) m2.ensure_stiffness();
        const size_t n = m2.level_n((int)l);
        P2Stage st{*m2.sim, where};
        double *du = st.inout(u, n, 0);
        const double *db = st.in(b, n, 1);
        m2.smooth((int)l, du, db, forward != 0);
        st.finish(u, du, n);
        return VF_OK;
    }
    mg->sync_masks();
    if (l >= 1) mg->ensure_blocks<double>();
    const size_t n = level_n(mg, l);
    Staged st{mg->sim, where};
    double *du = where == VF_DEVICE ? u : mg->sim->stage_out(n);
    if (where == VF_HOST) h2d(du, u, n, mg->st());
    const double *db = st.in(b, n, mg->sim->stage_aux(n));
    mg->smooth<double>((int)l, du, db, forward != 0);
    st.finish(u, du, n);
    VF_API_END
}

int vf_mg_restrict(vf_mg *mg, int64_t l, const double *fine, double *coarse, int where) {
    VF_API_BEGIN
    need(mg && fine && coarse, "null argument");
    need(l >= 0 && l < mg->L, "level out of range");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.sync_masks();
        const size_t nf = m2.level_n((int)l), nc = m2.level_n((int)l + 1);
        P2Stage st{*m2.sim, where};
        const double *df = st.in(fine, nf, 0);
        double *dc = st.out(coarse, nc, 1);
        m2.restrict_to((int)l, df, dc);
        st.finish(coarse, dc, nc);
        return VF_OK;
    }
    mg->sync_masks();
    const size_t nf = level_n(mg, l), nc = level_n(mg, l + 1);
    Staged st{mg->sim, where};
    const double *df = st.in(fine, nf, mg->sim->stage_in(nf));
    double *dc = st.out(coarse, mg->sim->stage_out(nc));
    launch_restrict<double>(mg->lv[l]->g, mg->lv[l + 1]->g, df, dc, mg->st());
    st.finish(coarse, dc, nc);
    VF_API_END
}

int vf_mg_interpolate(vf_mg *mg, int64_t l, const double *coarse, double *fine, int accumulate, int where) {
    VF_API_BEGIN
    need(mg && fine && coarse, "null argument");
    need(l >= 0 && l < mg->L, "level out of range");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.sync_masks();
        const size_t nf = m2.level_n((int)l), nc = m2.level_n((int)l + 1);
        P2Stage st{*m2.sim, where};
        const double *dc = st.in(coarse, nc, 0);
        double *df = st.inout(fine, nf, 1);
        m2.interpolate((int)l, dc, df, accumulate != 0);
        st.finish(fine, df, nf);
        return VF_OK;
    }
    mg->sync_masks();
    const size_t nf = level_n(mg, l), nc = level_n(mg, l + 1);
    Staged st{mg->sim, where};
    const double *dc = st.in(coarse, nc, mg->sim->stage_in(nc));
    double *df = where == VF_DEVICE ? fine : mg->sim->stage_out(nf);
    if (where == VF_HOST) h2d(df, fine, nf, mg->st());
    launch_interpolate<double>(mg->lv[l]->g, mg->lv[l + 1]->g, dc, df, accumulate, mg->st());
    st.finish(fine, df, nf);
    VF_API_END
}

int vf_mg_coarse_solve(vf_mg *mg, const double *b, double *x, int where) {
    VF_API_BEGIN
    need(mg && b && x, "null argument");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        m2.ensure_stiffness();
        const size_t n = m2.level_n(m2.L);
        P2Stage st{*m2.sim, where};
        const double *db = st.in(b, n, 0);
        double *dx = st.out(x, n, 1);
        m2.coarse_solve(db, dx);
        st.finish(x, dx, n);
        return VF_OK;
    }
    mg->sync_masks();
    mg->ensure_stiffness();
    const size_t n = level_n(mg, mg->L);
    Staged st{mg->sim, where};
    const double *db = st.in(b, n, mg->sim->stage_in(n));
    double *dx = st.out(x, mg->sim->stage_out(n));
    mg->coarse_solve<double>(db, dx);
    st.finish(x, dx, n);
    VF_API_END
}

int vf_mg_zero_dirichlet(vf_mg *mg, int64_t l, double *u, int where) {
    VF_API_BEGIN
    need(mg && u, "null argument");
    need(l >= 0 && l <= mg->L, "level out of range");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        const size_t n = m2.level_n((int)l);
        P2Stage st{*m2.sim, where};
        double *du = st.inout(u, n, 0);
        m2.zero_dirichlet((int)l, du);
        st.finish(u, du, n);
        return VF_OK;
    }
    const size_t n = level_n(mg, l);
    double *du = where == VF_DEVICE ? u : mg->sim->stage_out(n);
    if (where == VF_HOST) h2d(du, u, n, mg->st());
    mg->zero_dirichlet<double>((int)l, du);
    if (where == VF_HOST) { d2h(u, du, n, mg->st()); sync(mg->st()); }
    VF_API_END
}

int vf_mg_set_overlap_setup(vf_mg *mg, int enable) {
    VF_API_BEGIN
    need(mg, "null argument");
    mg->overlapSetup = enable != 0;
    VF_API_END
}

int vf_mg_set_partial_updates(vf_mg *mg, int enable) {
    VF_API_BEGIN
    need(mg, "null argument");
    mg->partialUpdates = enable != 0;
    VF_API_END
}

int vf_mg_partial_update_count(vf_mg *mg, int64_t *count) {
    VF_API_BEGIN
    need(mg && count, "null argument");
    *count = mg->partialCount;
    VF_API_END
}

int vf_mg_debug_get(vf_mg *mg, int64_t l, int which, double *out, int where) {
    VF_API_BEGIN
    need(mg && out, "null argument");
    need(l >= 0 && l <= mg->L, "level out of range");
    need(which == 0 || which == 1, "which must be 0 (x) or 1 (b)");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        const size_t n = m2.level_n((int)l);
        const double *src = which ? m2.lv[l]->b.data() : m2.lv[l]->x.data();
        if (where == VF_DEVICE) d2d(out, src, n, m2.st());
        else d2h(out, src, n, m2.st());
        sync(m2.st());
        return VF_OK;
    }
    if (mg->dbgPrec < 0) throw std::runtime_error("no multigrid cycle has run yet");
    const size_t n = level_n(mg, l);
    hipStream_t s = mg->st();
    double *dst = where == VF_DEVICE ? out : mg->sim->stage_out(n);
    if (mg->dbgPrec == VF_PREC_FP64) {
        const double *src = l == 0 ? (const double *)(which ? mg->dbgB0 : mg->dbgX0)
                                   : (which ? mg->lv[l]->b.data() : mg->lv[l]->x.data());
        need(src != nullptr && (l == 0 || (which ? mg->lv[l]->b.size() : mg->lv[l]->x.size()) >= n),
             "the last cycle did not use this level's buffer");
        d2d(dst, src, n, s);
    } else {
        const float *src = l == 0 ? (const float *)(which ? mg->dbgB0 : mg->dbgX0)
                                  : (which ? mg->lv[l]->b32.data() : mg->lv[l]->x32.data());
        need(src != nullptr && (l == 0 || (which ? mg->lv[l]->b32.size() : mg->lv[l]->x32.size()) >= n),
             "the last cycle did not use this level's buffer");
        launch_convert<double, float>(n, src, dst, s);
    }
    if (where == VF_HOST) d2h(out, dst, n, s);
    sync(s);
    VF_API_END
}

int vf_mg_solve(vf_mg *mg, double *x, const double *f, int64_t numSteps, int64_t smoothing, int stiffnessUpdated,
                int zeroDirichlet, int fullMultigrid, vf_mg_callback cb, void *user, int where) {
    VF_API_BEGIN
    need(mg && x && f, "null argument");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        const size_t n = m2.level_n(0);
        P2Stage st{*m2.sim, where};
        double *dx = st.inout(x, n, 0);
        const double *df = st.in(f, n, 1);
        m2.solve(dx, df, (int)numSteps, (int)smoothing, stiffnessUpdated != 0, zeroDirichlet != 0, fullMultigrid != 0,
                 cb, user);
        st.finish(x, dx, n);
        return VF_OK;
    }
    const size_t n = level_n(mg, 0);
    hipStream_t s = mg->st();
    double *dx = x;
    const double *df = f;
    if (where == VF_HOST) {
        mg->px.resize(n);
        mg->pb.resize(n);
        h2d(mg->px.data(), x, n, s);
        h2d(mg->pb.data(), f, n, s);
        dx = mg->px.data();
        df = mg->pb.data();
    }
    mg->solve(dx, df, (int)numSteps, (int)smoothing, stiffnessUpdated != 0, zeroDirichlet != 0, fullMultigrid != 0, cb,
              user);
    mg->dbg_keep_level0(n);
    if (where == VF_HOST) d2h(x, dx, n, s);
    sync(s);
    VF_API_END
}

int vf_mg_precondition(vf_mg *mg, const double *r, double *s, int64_t mgIterations, int64_t smoothing,
                       int fullMultigrid, int precision, int where) {
    VF_API_BEGIN
    need(mg && r && s, "null argument");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        need(precision == VF_PREC_FP64, "planar grids run the FP64 V-cycle only");
        const size_t n = m2.level_n(0);
        P2Stage st{*m2.sim, where};
        const double *dr = st.in(r, n, 0);
        double *ds = st.out(s, n, 1);
        m2.precondition(dr, ds, (int)mgIterations, (int)smoothing, fullMultigrid != 0);
        st.finish(s, ds, n);
        return VF_OK;
    }
    const size_t n = level_n(mg, 0);
    Staged st{mg->sim, where};
    // the preconditioner reads r while writing s: keep them in distinct staging buffers
    const double *dr = st.in(r, n, mg->sim->stage_in(n));
    double *ds = st.out(s, mg->sim->stage_out(n));
    mg->precondition(dr, ds, (int)mgIterations, (int)smoothing, fullMultigrid != 0, precision);
    mg->dbg_keep_level0(n);
    st.finish(s, ds, n);
    VF_API_END
}

int vf_mg_pcg(vf_mg *mg, double *x, const double *b, int64_t maxIter, double tol, int64_t mgIterations,
              int64_t smoothing, int fullMultigrid, int precision, int dirichletAlreadySatisfied, vf_pcg_callback cb,
              void *user, int where, vf_pcg_stats *stats) {
    VF_API_BEGIN
    need(mg && x && b, "null argument");
    need(precision == VF_PREC_FP64 || precision == VF_PREC_FP32, "unknown precision");
    if (mg->p2) {
        auto &m2 = *mg->p2;
        need(precision == VF_PREC_FP64, "planar grids run the FP64 V-cycle only");
        const size_t n = m2.level_n(0);
        P2Stage st{*m2.sim, where};
        double *dx = st.inout(x, n, 0);
        const double *db = st.in(b, n, 1);
        const auto t0 = std::chrono::steady_clock::now();
        double rn = 0, bn = 0;
        const int64_t its = m2.pcg(dx, db, maxIter, tol, (int)mgIterations, (int)smoothing, fullMultigrid != 0,
                                   dirichletAlreadySatisfied != 0, cb, user, &rn, &bn);
        st.finish(x, dx, n);
        if (stats) {
            stats->iterations = its;
            stats->residual_norm = rn;
            stats->rhs_norm = bn;
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return VF_OK;
    }
    const size_t n = level_n(mg, 0);
    hipStream_t s = mg->st();
    double *dx = x;
    const double *db = b;
    if (where == VF_HOST) {
        mg->px.resize(n);
        mg->pb.resize(n);
        h2d(mg->px.data(), x, n, s);
        h2d(mg->pb.data(), b, n, s);
        dx = mg->px.data();
        db = mg->pb.data();
    }
    mg->pcg(dx, db, maxIter, tol, (int)mgIterations, (int)smoothing, fullMultigrid != 0, precision,
            dirichletAlreadySatisfied != 0, cb, user, stats);
    if (where == VF_HOST) d2h(x, dx, n, s);
    sync(s);
    VF_API_END
}

// c = A^+ b for a symmetric s x s matrix A (row-major): eigen-decomposition by cyclic Jacobi rotations, then
// the pseudo-inverse that drops the singular values |lambda| below s * eps * max |lambda| (the default
// rank threshold of the Jacobi SVD solve the reference uses, LayerByLayer.hh:120-121).
static std::vector<double> sym_pinv_solve(std::vector<double> A, const std::vector<double> &b, int s) {
    std::vector<double> V((size_t)s * s, 0.0);
    for (int i = 0; i < s; ++i) V[(size_t)i * s + i] = 1.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0, diag = 0;
        for (int p = 0; p < s; ++p) {
            diag += A[(size_t)p * s + p] * A[(size_t)p * s + p];
            for (int q = p + 1; q < s; ++q) off += A[(size_t)p * s + q] * A[(size_t)p * s + q];
        }
        if (off <= 1e-32 * diag || off == 0) break;
        for (int p = 0; p < s; ++p)
            for (int q = p + 1; q < s; ++q) {
                const double apq = A[(size_t)p * s + q];
                if (apq == 0) continue;
                const double theta = (A[(size_t)q * s + q] - A[(size_t)p * s + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < s; ++k) {  // columns p, q
                    const double akp = A[(size_t)k * s + p], akq = A[(size_t)k * s + q];
                    A[(size_t)k * s + p] = c * akp - sn * akq;
                    A[(size_t)k * s + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < s; ++k) {  // rows p, q
                    const double apk = A[(size_t)p * s + k], aqk = A[(size_t)q * s + k];
                    A[(size_t)p * s + k] = c * apk - sn * aqk;
                    A[(size_t)q * s + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < s; ++k) {
                    const double vkp = V[(size_t)k * s + p], vkq = V[(size_t)k * s + q];
                    V[(size_t)k * s + p] = c * vkp - sn * vkq;
                    V[(size_t)k * s + q] = sn * vkp + c * vkq;
                }
            }
    }
    double smax = 0;
    for (int k = 0; k < s; ++k) smax = std::max(smax, std::fabs(A[(size_t)k * s + k]));
    const double thr = std::max(smax * s * std::numeric_limits<double>::epsilon(), std::numeric_limits<double>::min());
    std::vector<double> c(s, 0.0);
    for (int k = 0; k < s; ++k) {
        const double lam = A[(size_t)k * s + k];
        if (!(std::fabs(lam) >= thr)) continue;
        double vb = 0;
        for (int i = 0; i < s; ++i) vb += V[(size_t)i * s + k] * b[i];
        for (int i = 0; i < s; ++i) c[i] += V[(size_t)i * s + k] * vb / lam;
    }
    return c;
}

// The layer-by-layer evaluation on a planar grid: the same steps as lbl_run below (LayerByLayer.hh:223-300 with
// the guesses of LayerByLayer.hh:24-217) on the planar simulator and hierarchy.
static void lbl_run_2d(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations,
                       int64_t smoothing, int fullMultigrid, int precision, int method, int64_t history, bool zeroInit,
                       double *uFull, bool uFullValid, int uFullWhere, double *objective, double *grad, int gradWhere,
                       double *perLayerCompliance, int64_t *perLayerIterations, vf_pcg_callback itcb, void *ituser,
                       vf_lbl_callback lblcb, void *lbluser) {
    need(precision == VF_PREC_FP64, "planar grids run the FP64 V-cycle only");
    vf::planar::PMG &m = *mg->p2;
    vf::planar::PSim &sim = *m.sim;
    hipStream_t s = sim.stream;
    const int64_t numLayers = sim.ne[1];
    const size_t n = sim.ndof(), ne = sim.nel();
    const int H = method == VF_LBL_INIT_ZERO ? 0 : (int)history;
    DBuf<double> f(n), u(n), g(ne);
    std::vector<DBuf<double>> hist(H), Ku(method == VF_LBL_INIT_SUBSPACE ? H : 0);
    for (auto &h : hist) h.resize(n);
    for (auto &k : Ku) k.resize(n);
    std::deque<int> order;  // history slots, newest first
    launch_fill<double>(ne, g.data(), 0.0, s);
    sim.set_mask_height(sim.h[1] * (double)numLayers, true);
    double total = 0;
    int64_t count = 0;
    for (int64_t l = numLayers; l > 0; l -= std::min(layerIncrement, l)) {
        if (l < numLayers) sim.remove_top_layers(layerIncrement);
        m.sync_masks();
        sim.build_load(f.data());
        const int hs = (int)order.size();
        if (l == numLayers && !zeroInit && uFull && uFullValid) {
            if (uFullWhere == VF_DEVICE) d2d(u.data(), uFull, n, s);
            else h2d(u.data(), uFull, n, s);
        } else if (hs == 0) {
            launch_fill<double>(n, u.data(), 0.0, s);
        } else if (method == VF_LBL_INIT_FD) {
            d2d(u.data(), hist[order[0]].data(), n, s);
            if (hs >= 2) {  // 2 u_0 - u_1
                launch_axpy_plain(n, 1.0, hist[order[0]].data(), u.data(), s);
                launch_axpy_plain(n, -1.0, hist[order[1]].data(), u.data(), s);
            }
        } else {  // subspace
            for (int j = 0; j < hs; ++j) sim.apply_K(hist[order[j]].data(), Ku[j].data());
            std::vector<double> A((size_t)hs * hs), bb(hs);
            for (int i = 0; i < hs; ++i) {
                bb[i] = sim.dot(hist[order[i]].data(), f.data(), true);
                for (int j = 0; j <= i; ++j)
                    A[(size_t)i * hs + j] = A[(size_t)j * hs + i] = sim.dot(hist[order[i]].data(), Ku[j].data(), true);
            }
            const std::vector<double> c = sym_pinv_solve(A, bb, hs);
            launch_fill<double>(n, u.data(), 0.0, s);
            for (int i = 0; i < hs; ++i) launch_axpy_plain(n, c[i], hist[order[i]].data(), u.data(), s);
        }
        sim.zero_detached(u.data());
        const int64_t its = m.pcg(u.data(), f.data(), maxIter, tol, (int)mgIterations, (int)smoothing,
                                  fullMultigrid != 0, true, itcb, ituser, nullptr, nullptr);
        const double comp = sim.dot(f.data(), u.data(), true);
        if (lblcb) {
            DBuf<double> gl(ne);
            sim.compliance_gradient(u.data(), gl.data(), false);
            std::vector<double> hgl(ne), hu(n);
            d2h(hgl.data(), gl.data(), ne, s);
            d2h(hu.data(), u.data(), n, s);
            if (lblcb(lbluser, l, comp, hgl.data(), hu.data())) throw std::runtime_error("layer-by-layer run aborted by callback");
        }
        if (perLayerCompliance) perLayerCompliance[count] = comp;
        if (perLayerIterations) perLayerIterations[count] = its;
        total += comp;
        sim.compliance_gradient(u.data(), g.data(), true);
        ++count;
        if (l == numLayers && uFull) {
            if (uFullWhere == VF_DEVICE) d2d(uFull, u.data(), n, s);
            else d2h(uFull, u.data(), n, s);
        }
        if (H > 0 && l >= layerIncrement) {
            int slot;
            if ((int)order.size() < H) {
                slot = (int)order.size();
            } else {
                slot = order.back();
                order.pop_back();
            }
            d2d(hist[slot].data(), u.data(), n, s);
            order.push_front(slot);
        }
    }
    *objective = 0.5 * total / (double)count;
    std::vector<double> hg(ne);
    d2h(hg.data(), g.data(), ne, s);
    for (auto &v : hg) v /= (double)count;
    if (gradWhere == VF_DEVICE) h2d(grad, hg.data(), ne, s);
    else std::copy(hg.begin(), hg.end(), grad);
    sync(s);
}

// LayerByLayerEvaluator::run (LayerByLayer.hh:223-300) with its initial guesses (LayerByLayer.hh:24-217):
// zero; the previous layers' solutions extrapolated (history 1: copied, history 2: 2 u_0 - u_1); or the
// Galerkin projection of the next system onto the last N solutions, g = sum_i c_i u_i with
// A c = b, A_ij = u_i . K u_j and b_i = u_i . f over the non-detached nodes (the reference updates A and b
// by recurrences that equal these products, LayerByLayer.hh:140-208). Guesses are zero on detached rows
// and keep the solutions' zero Dirichlet components.
static void lbl_run(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations,
                    int64_t smoothing, int fullMultigrid, int precision, int method, int64_t history, bool zeroInit,
                    double *uFull, bool uFullValid, int uFullWhere, double *objective, double *grad, int gradWhere,
                    double *perLayerCompliance, int64_t *perLayerIterations, vf_pcg_callback itcb = nullptr,
                    void *ituser = nullptr, vf_lbl_callback lblcb = nullptr, void *lbluser = nullptr) {
    if (mg->p2) {
        lbl_run_2d(mg, layerIncrement, maxIter, tol, mgIterations, smoothing, fullMultigrid, precision, method, history,
                   zeroInit, uFull, uFullValid, uFullWhere, objective, grad, gradWhere, perLayerCompliance,
                   perLayerIterations, itcb, ituser, lblcb, lbluser);
        return;
    }
    vf_sim *sim = mg->sim;
    hipStream_t s = mg->st();
    const int64_t numLayers = sim->ne[1];
    const size_t n = 3 * sim->nn(), ne = sim->nel();
    const int H = method == VF_LBL_INIT_ZERO ? 0 : (int)history;
    DBuf<double> f(n), u(n), g(ne), work(n);
    std::vector<DBuf<double>> hist(H), Ku(method == VF_LBL_INIT_SUBSPACE ? H : 0);
    for (auto &h : hist) h.resize(n);
    for (auto &k : Ku) k.resize(n);
    std::deque<int> order;  // history slots, newest first
    launch_fill<double>(ne, g.data(), 0.0, s);
    sim->set_mask_height(sim->h[1] * (double)numLayers, true);
    double *sc = sim->scalars();
    auto dot = [&](const double *a, const double *b2) {
        launch_dot_active(sim->g, a, b2, sim->red_partials(reduction_partials(sim->g)), sc + 10, s);
        return sim->read_scalar(sc + 10);
    };
    double total = 0;
    int64_t count = 0;
    for (int64_t l = numLayers; l > 0; l -= std::min(layerIncrement, l)) {
        if (l < numLayers) sim->remove_top_layers(layerIncrement);
        mg->sync_masks();
        if (sim->hasForce) d2d(f.data(), sim->force.data(), n, s);
        else launch_fill<double>(n, f.data(), 0.0, s);
        if (sim->grav[0] != 0 || sim->grav[1] != 0 || sim->grav[2] != 0)
            launch_self_weight(sim->g, sim->rho.data(), sim->grav, sim->h[0] * sim->h[1] * sim->h[2], f.data(), s);
        // initial guess (the full design may start from the previous run's full-design solution)
        const int hs = (int)order.size();
        if (l == numLayers && !zeroInit && uFull && uFullValid) {
            if (uFullWhere == VF_DEVICE) d2d(u.data(), uFull, n, s);
            else h2d(u.data(), uFull, n, s);
        } else if (hs == 0) {
            launch_fill<double>(n, u.data(), 0.0, s);
        } else if (method == VF_LBL_INIT_FD) {
            d2d(u.data(), hist[order[0]].data(), n, s);
            if (hs >= 2) {  // 2 u_0 - u_1
                launch_axpy_plain(n, 1.0, hist[order[0]].data(), u.data(), s);
                launch_axpy_plain(n, -1.0, hist[order[1]].data(), u.data(), s);
            }
        } else {  // subspace
            for (int j = 0; j < hs; ++j) sim_apply_K<double>(sim, hist[order[j]].data(), Ku[j].data(), s);
            std::vector<double> A((size_t)hs * hs), bb(hs);
            for (int i = 0; i < hs; ++i) {
                bb[i] = dot(hist[order[i]].data(), f.data());
                for (int j = 0; j <= i; ++j) A[(size_t)i * hs + j] = A[(size_t)j * hs + i] = dot(hist[order[i]].data(), Ku[j].data());
            }
            const std::vector<double> c = sym_pinv_solve(A, bb, hs);
            launch_fill<double>(n, u.data(), 0.0, s);
            for (int i = 0; i < hs; ++i) launch_axpy_plain(n, c[i], hist[order[i]].data(), u.data(), s);
        }
        launch_zero_detached<double>(sim->g, u.data(), s);
        const int64_t its = mg->pcg(u.data(), f.data(), maxIter, tol, (int)mgIterations, (int)smoothing,
                                    fullMultigrid != 0, precision, true, itcb, ituser, nullptr);
        const double comp = dot(f.data(), u.data());
        if (lblcb) {
            // lblCallback(l, compliance, complianceGradientFlattened(u), u) (LayerByLayer.hh:277-279)
            DBuf<double> gl(ne);
            launch_compliance_gradient(sim->g, sim->K0.data(), sim->rho.data(), u.data(), gl.data(), sim->law, sim->E0,
                                       sim->Emin, sim->gamma, sim->q, sim->grav, sim->h[0] * sim->h[1] * sim->h[2], 0, s);
            std::vector<double> hgl(ne), hu(n);
            d2h(hgl.data(), gl.data(), ne, s);
            d2h(hu.data(), u.data(), n, s);
            sync(s);
            if (lblcb(lbluser, l, comp, hgl.data(), hu.data())) throw std::runtime_error("layer-by-layer run aborted by callback");
        }
        if (perLayerCompliance) perLayerCompliance[count] = comp;
        if (perLayerIterations) perLayerIterations[count] = its;
        total += comp;
        launch_compliance_gradient(sim->g, sim->K0.data(), sim->rho.data(), u.data(), g.data(), sim->law, sim->E0,
                                   sim->Emin, sim->gamma, sim->q, sim->grav, sim->h[0] * sim->h[1] * sim->h[2], 1, s);
        ++count;
        if (l == numLayers && uFull) {
            if (uFullWhere == VF_DEVICE) d2d(uFull, u.data(), n, s);
            else { d2h(uFull, u.data(), n, s); sync(s); }
        }
        // finalizeLayer: the solution joins the history (the oldest slot is reused once the history is full)
        if (H > 0 && l >= layerIncrement) {
            int slot;
            if ((int)order.size() < H) {
                slot = (int)order.size();
            } else {
                slot = order.back();
                order.pop_back();
            }
            d2d(hist[slot].data(), u.data(), n, s);
            order.push_front(slot);
        }
    }
    *objective = 0.5 * total / (double)count;
    std::vector<double> hg(ne);
    d2h(hg.data(), g.data(), ne, s);
    sync(s);
    for (auto &v : hg) v /= (double)count;
    if (gradWhere == VF_DEVICE) { h2d(grad, hg.data(), ne, s); sync(s); }
    else std::copy(hg.begin(), hg.end(), grad);
}

int vf_lbl_run_zero(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations,
                    int64_t smoothing, int fullMultigrid, int precision, double *objective, double *grad, int gradWhere,
                    double *perLayerCompliance, int64_t *perLayerIterations) {
    VF_API_BEGIN
    need(mg && objective && grad, "null argument");
    need(layerIncrement >= 1, "layerIncrement must be positive");
    lbl_run(mg, layerIncrement, maxIter, tol, mgIterations, smoothing, fullMultigrid, precision, VF_LBL_INIT_ZERO, 0,
            true, nullptr, false, VF_HOST, objective, grad, gradWhere, perLayerCompliance, perLayerIterations);
    VF_API_END
}

int vf_lbl_run(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations, int64_t smoothing,
               int fullMultigrid, int precision, int initMethod, int64_t history, int zeroInit, double *uFull,
               int uFullValid, int uFullWhere, double *objective, double *grad, int gradWhere,
               double *perLayerCompliance, int64_t *perLayerIterations, vf_pcg_callback itcb, void *ituser,
               vf_lbl_callback lblcb, void *lbluser) {
    VF_API_BEGIN
    need(mg && objective && grad, "null argument");
    need(layerIncrement >= 1, "layerIncrement must be positive");
    need(initMethod == VF_LBL_INIT_ZERO || initMethod == VF_LBL_INIT_FD || initMethod == VF_LBL_INIT_SUBSPACE,
         "initMethod must be VF_LBL_INIT_ZERO, VF_LBL_INIT_FD or VF_LBL_INIT_SUBSPACE");
    need(initMethod != VF_LBL_INIT_FD || history == 1 || history == 2, "the extrapolating guess keeps 1 or 2 solutions");
    need(initMethod != VF_LBL_INIT_SUBSPACE || (history >= 1 && history <= 64), "subspace history must be 1..64");
    lbl_run(mg, layerIncrement, maxIter, tol, mgIterations, smoothing, fullMultigrid, precision, initMethod, history,
            zeroInit != 0, uFull, uFullValid != 0, uFullWhere, objective, grad, gradWhere, perLayerCompliance,
            perLayerIterations, itcb, ituser, lblcb, lbluser);
    VF_API_END
}

// ---------------------------------------------------------------- stacked layer-by-layer batches
// The zero-initial-guess layer-by-layer evaluation (LayerByLayer.hh:223-296) solves one partial structure
// per layer. A batch of them is solved here as one problem: the structures lie side by side along x in
// segments of P = W + 2^L element columns (W the design's width), separated by 2^L columns of void, one
// element of the coarsest level. Every level's operator is then block-diagonal over the segments (the
// void columns couple nothing, their interior nodes and each segment's detached rows are constrained), so
// one hierarchy, one Galerkin chain and one V-cycle launch sequence serve the whole batch and reproduce,
// segment by segment, the V-cycle of each structure on its own. The outer PCG keeps one set of scalars per
// segment and stops each segment at its own tolerance.
namespace {
constexpr int kSegActive = 0, kSegFinishing = 1, kSegDone = 2;
}  // namespace

// A stacked batch: its grid, masks, hierarchy and work buffers (kept between evaluations with the same
// layers, constraints and element size; everything numeric is refreshed each time)
struct StackedBatch {
    std::vector<int> top;  // first masked element row of each segment's structure, tallest first
    int rows = 0, P = 0, L = 0;
    uint64_t bcVersion = 0;
    double h[3] = {0, 0, 0};
    std::unique_ptr<vf_sim> sim;
    std::unique_ptr<vf_mg> mg;
    DBuf<int> dtop, dstate;
    DBuf<double> f, x, sc, part, gst, xx;
};

namespace {
// the stacked simulator's geometry and the design's constraints in every segment (their values are not
// needed: the zero-guess evaluation does not impose them); the void columns and each segment's detached
// rows are held through the per-node masks of every level (constrain_stacked)
std::unique_ptr<vf_sim> make_stacked_sim(const vf_sim &base, const std::vector<int> &top, int P, int rows) {
    const int nseg = (int)top.size();
    auto c = std::make_unique<vf_sim>();
    c->stream = base.stream;
    const int64_t ne[3] = {(int64_t)nseg * P - (P - base.ne[0]), rows, base.ne[2]};
    for (int d = 0; d < 3; ++d) {
        c->ne[d] = ne[d];
        c->h[d] = base.h[d];
        c->lo[d] = base.lo[d];
        c->hi[d] = base.lo[d] + base.h[d] * (double)ne[d];
    }
    if ((ne[0] + 1) * (ne[1] + 1) * (ne[2] + 1) * 3 >= INT_MAX) throw std::runtime_error("stacked grid too large");
    c->g = make_grid(ne);
    hipStream_t s = c->stream;
    c->rho.resize(c->nel());
    c->E.resize(c->nel());
    c->E32.resize(c->nel());
    const size_t nn = c->nn();
    const Grid &g = c->g, &gb = base.g;
    c->hmask.assign(nn, 0);
    std::vector<size_t> held;  // the design's constrained nodes below the stacked grid's top
    for (size_t m = 0; m < base.nn(); ++m)
        if (base.hmask[m] && (int)((m / gb.nn[2]) % gb.nn[1]) < g.nn[1]) held.push_back(m);
    std::vector<int> nodes;
    std::vector<uint8_t> cm;
    nodes.reserve(held.size() * nseg);
    cm.reserve(held.size() * nseg);
    for (int k = 0; k < nseg; ++k)
        for (size_t m : held) {
            const uint8_t v = base.hmask[m];
            const int lx = (int)(m / ((size_t)gb.nn[1] * gb.nn[2])), iy = (int)((m / gb.nn[2]) % gb.nn[1]),
                      iz = (int)(m % gb.nn[2]);
            const size_t n = ((size_t)(k * P + lx) * g.nn[1] + iy) * g.nn[2] + iz;
            c->hmask[n] = v;
            nodes.push_back((int)n);
            cm.push_back(v);
        }
    c->ndir = (int)nodes.size();
    c->dmask.resize(nn);
    h2d(c->dmask.data(), c->hmask.data(), nn, s);
    c->dnodes.resize(std::max<size_t>(nodes.size(), 1));
    c->dnmask.resize(std::max<size_t>(cm.size(), 1));
    h2d(c->dnodes.data(), nodes.data(), nodes.size(), s);
    h2d(c->dnmask.data(), cm.data(), cm.size(), s);
    c->dvals.resize(3 * std::max<size_t>(nodes.size(), 1));
    launch_fill<double>(c->dvals.size(), c->dvals.data(), 0.0, s);
    // every segment's structure is below the tallest one's mask height
    c->maskHeight = base.h[1] * (double)top[0];
    c->firstMasked = top[0];
    c->firstDetached = top[0] + 1;
    c->refresh_grid();
    sync(s);
    return c;
}

// the numeric state of the stacked simulator from the design's: material, law, element tables, moduli,
// densities, forces and gravity
void refresh_stacked_sim(vf_sim &c, const vf_sim &base, int P, const int *dtop, int nseg, const double *Efull,
                         const float *E32full) {
    c.young = base.young; c.poisson = base.poisson; c.law = base.law;
    c.E0 = base.E0; c.Emin = base.Emin; c.gamma = base.gamma; c.q = base.q;
    std::copy(base.grav, base.grav + 3, c.grav);
    std::copy(base.K0h, base.K0h + 576, c.K0h);
    c.structured = base.structured;
    c.walsh = base.walsh;
    hipStream_t s = c.stream;
    auto dup = [&](auto &dst, const auto &src) {
        dst.resize(src.size());
        d2d(dst.data(), src.data(), src.size(), s);
    };
    dup(c.K0, base.K0); dup(c.K032, base.K032); dup(c.Ftab, base.Ftab); dup(c.Ftab32, base.Ftab32);
    dup(c.Wtab, base.Wtab); dup(c.Wtab32, base.Wtab32);
    launch_stack_elems<double>(c.g, base.g, P, nseg, dtop, base.rho.data(), c.rho.data(), s);
    launch_stack_elems<double>(c.g, base.g, P, nseg, dtop, Efull, c.E.data(), s);
    launch_stack_elems<float>(c.g, base.g, P, nseg, dtop, E32full, c.E32.data(), s);
    c.hasForce = base.hasForce;
    if (base.hasForce) {
        c.force.resize(3 * c.nn());
        launch_stack_nodes(c.g, base.g, P, nseg, dtop, base.force.data(), c.force.data(), s);
    }
    ++c.version;
    ++c.baseVersion;
}
}  // namespace

// first detached node row of each segment's structure on every level (vf_mg::sync_masks for a structure of
// that height): ydet[l][k]
static std::vector<std::vector<int>> stacked_ydet(const vf_mg &m, const std::vector<int> &top, double hy) {
    std::vector<std::vector<int>> y(m.L + 1, std::vector<int>(top.size()));
    for (int l = 0; l <= m.L; ++l)
        for (size_t k = 0; k < top.size(); ++k) {
            const Grid &g = m.lv[l]->g;
            if (l == 0) {
                y[l][k] = std::min(top[k] + 1, g.nn[1]);
            } else {
                const int64_t fm = (int64_t)std::ceil(hy * (double)top[k] / m.lv[l]->h[1] - 1e-10);
                y[l][k] = (int)std::min<int64_t>(fm + 1, g.nn[1]);
            }
        }
    return y;
}

// hold, on every level, the nodes of the void columns and each segment's rows from its first detached row
// on: their per-node masks become 7, so the smoothers skip them and the products and residuals are zero
// there; the constraint lists keep the design's constraints only
static void constrain_stacked(vf_mg &m, const std::vector<std::vector<int>> &ydet, int P) {
    const int nseg = (int)ydet[0].size();
    DBuf<int> dy(nseg);
    for (int l = 0; l <= m.L; ++l) {
        const Grid &g = m.lv[l]->g;
        h2d(dy.data(), ydet[l].data(), nseg, m.st());
        uint8_t *mask = l == 0 ? m.sim->dmask.data() : m.lv[l]->dmask.data();
        launch_stack_mask(g, P >> l, (P - (1 << m.L)) >> l, nseg, dy.data(), mask, m.st());
    }
    sync(m.st());
}

// the batch for these layers from the solver's cache, or a new one
static StackedBatch &stacked_batch_for(vf_mg *base, std::vector<std::shared_ptr<StackedBatch>> &next,
                                       const std::vector<int> &top, int rows) {
    vf_sim *bs = base->sim;
    const int L = base->L, P = (int)bs->ne[0] + (1 << L);
    for (auto &b : base->stackCache)
        if (b && b->top == top && b->rows == rows && b->P == P && b->L == L && b->bcVersion == bs->bcVersion &&
            std::equal(b->h, b->h + 3, bs->h)) {
            next.push_back(b);
            b.reset();
            return *next.back();
        }
    auto b = std::make_shared<StackedBatch>();
    const int nseg = (int)top.size(), W = (int)bs->ne[0];
    b->top = top;
    b->rows = rows;
    b->P = P;
    b->L = L;
    b->bcVersion = bs->bcVersion;
    std::copy(bs->h, bs->h + 3, b->h);
    hipStream_t s = bs->stream;
    b->dtop.resize(nseg);
    h2d(b->dtop.data(), top.data(), top.size(), s);
    b->sim = make_stacked_sim(*bs, top, P, rows);
    std::copy(bs->K0h, bs->K0h + 576, b->sim->K0h);
    b->mg = std::make_unique<vf_mg>(b->sim.get(), L);
    vf_mg &m = *b->mg;
    // the coarse levels take the design solver's constraint masks (coarsened when that solver was built,
    // MultigridSolver.hh:58-103), replicated into every segment, so each segment's hierarchy is that
    // solver's own
    for (int l = 1; l <= L; ++l) {
        LevelData &c = *m.lv[l];
        const LevelData &bl = *base->lv[l];
        const Grid &gc = c.g, &gb = bl.g;
        std::fill(c.hmask.begin(), c.hmask.end(), uint8_t(0));
        const int Pl = P >> l;
        for (size_t mm = 0; mm < bl.hmask.size(); ++mm) {
            const uint8_t v = bl.hmask[mm];
            if (!v) continue;
            const int lx = (int)(mm / ((size_t)gb.nn[1] * gb.nn[2])), iy = (int)((mm / gb.nn[2]) % gb.nn[1]),
                      iz = (int)(mm % gb.nn[2]);
            if (iy >= gc.nn[1]) continue;
            for (int k = 0; k < nseg; ++k) c.hmask[((size_t)(k * Pl + lx) * gc.nn[1] + iy) * gc.nn[2] + iz] = v;
        }
        upload_dirichlet(c.hmask, c.dmask, c.dnodes, c.dnmask, c.ndir, m.st());
    }
    m.partialUpdates = false;
    m.overlapSetup = false;
    const std::vector<std::vector<int>> ydet = stacked_ydet(m, top, bs->h[1]);
    constrain_stacked(m, ydet, P);
    m.seg = std::make_unique<vf_mg::SegCoarse>();
    m.seg->nseg = nseg;
    m.seg->Pc = P >> L;
    m.seg->Wc = W >> L;
    m.seg->ydet = ydet[L];
    const Grid &g = b->sim->g;
    const size_t n0 = 3 * (size_t)g.nnodes;
    b->f.resize(n0);
    b->x.resize(n0);
    b->sc.resize((size_t)16 * nseg);
    b->dstate.resize(nseg);
    b->gst.resize(b->sim->nel());
    b->xx.resize(1);
    const int xc = walsh_xchunk(g, AM_SET, P);
    const int tiles = ((g.ydet + 14) / 15) * ((g.nn[2] + 14) / 15), nch = (g.nn[0] + xc - 1) / xc;
    b->part.resize(std::max<size_t>(2 * (size_t)active_partials(g), (size_t)tiles * nch) + 64);
    m.pr.resize(n0); m.ps.resize(n0); m.pd.resize(n0); m.pAd.resize(n0);
    next.push_back(b);
    return *next.back();
}

// One stacked batch: heights top[k] (element rows, tallest first); adds each segment's compliance <f, u>
// and PCG iteration count to comp / its and its gradient (in segment order) to gsum.
static void lbl_stacked_batch(vf_mg *base, std::vector<std::shared_ptr<StackedBatch>> &next,
                              const std::vector<int> &top, int rows, int64_t maxIter, double tol, int mgIter, int nu,
                              bool fullMG, int prec, bool useGraphs, const double *Efull, const float *E32full,
                              double *comp, int64_t *its, double *gsum) {
    vf_sim *bs = base->sim;
    hipStream_t s = bs->stream;
    const int nseg = (int)top.size();
    StackedBatch &B = stacked_batch_for(base, next, top, rows);
    const int P = B.P;
    vf_sim *S = B.sim.get();
    refresh_stacked_sim(*S, *bs, P, B.dtop.data(), nseg, Efull, E32full);
    vf_mg &m = *B.mg;
    m.seg->state = B.dstate.data();
    m.copy_coarsening(*base);
    m.symmetricGS = base->symmetricGS;
    m.useGraphs = useGraphs;
    // the batch's own kernel choices: its solver times the kernel families on the stacked grid at its first
    // cycle in each precision (kept with the cached batch)
    m.sync_masks();
    const Grid &g = S->g;
    const size_t n0 = 3 * (size_t)g.nnodes;
    // load: the design's forces in every segment plus each structure's self-weight
    DBuf<double> &f = B.f, &x = B.x, &sc = B.sc, &part = B.part, &xx = B.xx;
    DBuf<int> &dstate = B.dstate;
    if (S->hasForce) d2d(f.data(), S->force.data(), n0, s);
    else launch_fill<double>(n0, f.data(), 0.0, s);
    if (S->grav[0] != 0 || S->grav[1] != 0 || S->grav[2] != 0)
        launch_self_weight(g, S->rho.data(), S->grav, S->h[0] * S->h[1] * S->h[2], f.data(), s);
    // zero initial guess; like the sequential evaluator, the constrained values are not imposed on it
    launch_fill<double>(n0, x.data(), 0.0, s);
    // PCG (vf_mg::pcg, fused form) with per-segment scalars sc[16 k ..]
    double *r = m.pr.data(), *sv = m.ps.data(), *d = m.pd.data(), *Ad = m.pAd.data();
    launch_fill<double>((size_t)16 * nseg, sc.data(), 0.0, s);
    const int xc = walsh_xchunk(g, AM_SET, P);
    const int tiles = ((g.ydet + 14) / 15) * ((g.nn[2] + 14) / 15), nch = (g.nn[0] + xc - 1) / xc;
    if (part.size() < (size_t)tiles * nch) throw std::runtime_error("stacked batch: partial-sum buffer too small");
    std::vector<int> state(nseg, kSegActive);
    std::vector<double> hsc((size_t)16 * nseg);
    auto read_sc = [&] {
        d2h(hsc.data(), sc.data(), hsc.size(), s);
        sync(s);
    };
    launch_seg_dot(g, f.data(), f.data(), part.data(), P, nseg, sc.data(), 7, s);
    launch_dot_active(g, x.data(), x.data(), part.data(), xx.data(), s);
    double hx = 0;
    d2h(&hx, xx.data(), 1, s);
    sync(s);
    launch_fill<double>(n0, r, 0.0, s);
    if (hx == 0.0) {
        launch_copy_active<double, double>(g, 0, f.data(), r, s);
        m.zero_dirichlet<double>(0, r);
        launch_zero_held<double>(g, S->dmask.data(), r, s);
    } else {
        m.apply<double>(0, x.data(), r, f.data(), AM_RESIDUAL, true);
    }
    launch_seg_dot(g, r, r, part.data(), P, nseg, sc.data(), 4, s);
    read_sc();
    const double tol2 = tol * tol;
    for (int k = 0; k < nseg; ++k) {
        const double rr = hsc[16 * k + 4], bb = hsc[16 * k + 7];
        if (std::isnan(rr)) throw NanError("NaN encountered");
        its[k] = 0;
        if (!(maxIter > 0 && rr > tol2 * bb)) state[k] = kSegDone;
    }
    h2d(dstate.data(), state.data(), state.size(), s);
    const bool flexible = prec == VF_PREC_FP32;
    // the element products read d everywhere, also in the void columns (times zero moduli) and in segments
    // that never iterate: it must hold finite values
    launch_fill<double>(n0, d, 0.0, s);
    float *r32 = nullptr;
    if (prec == VF_PREC_FP32) {
        m.ensure_buffers<float>(0);
        r32 = m.lv[0]->b32.data();
        launch_convert<float, double>(n0, r, r32, s);
    }
    auto body = [&](bool firstIt) {
        m.precondition(r, sv, mgIter, nu, fullMG, prec, true);
        const double *adp = (flexible && !firstIt) ? Ad : nullptr;
        launch_seg_scalars(sc.data(), dstate.data(), nseg, 0, s);
        if (prec == VF_PREC_FP32) {
            float *s32 = m.lv[0]->x32.data();
            m.zero_dirichlet<float>(0, s32);
            launch_zero_held<float>(g, S->dmask.data(), s32, s);
            launch_seg_dots_s<float>(g, r, s32, adp, part.data(), P, nseg, sc.data(), s);
            launch_seg_direction_x<float>(g, s32, d, x.data(), sc.data(), dstate.data(), P, nseg, flexible ? 1 : 0,
                                          firstIt ? 1 : 0, s);
        } else {
            m.zero_dirichlet<double>(0, sv);
            launch_zero_held<double>(g, S->dmask.data(), sv, s);
            launch_seg_dots_s<double>(g, r, sv, adp, part.data(), P, nseg, sc.data(), s);
            launch_seg_direction_x<double>(g, sv, d, x.data(), sc.data(), dstate.data(), P, nseg, 0, firstIt ? 1 : 0,
                                           s);
        }
        launch_fine_apply_walsh<double>(g, S->Wtab.data(), S->E.data(), d, Ad, nullptr, m.dm(0), AM_SET, s, part.data(),
                                        xx.data(), P);
        launch_seg_sum_chunks(part.data(), tiles, xc, nch, P, nseg, sc.data(), 3, s);
        launch_seg_scalars(sc.data(), dstate.data(), nseg, 1, s);
        launch_seg_residual(g, r, Ad, sc.data(), dstate.data(), P, nseg, part.data(), r32, s);
    };
    GraphGuard graph;
    auto capture = [&] {
        hipGraph_t gr = nullptr;
        VF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            body(false);
        } catch (...) {
            (void)hipStreamEndCapture(s, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        VF_HIP(hipStreamEndCapture(s, &gr));
        graph.graph = gr;
        VF_HIP(hipGraphInstantiate(&graph.exec, gr, nullptr, nullptr, 0));
    };
    int64_t it = 0;
    bool first = true, any = std::any_of(state.begin(), state.end(), [](int v) { return v == kSegActive; });
    if (any) m.ensure_stiffness();
    while (any) {
        ++it;
        if (useGraphs && !first && s != nullptr) {
            if (!graph.exec) capture();
            VF_HIP(hipGraphLaunch(graph.exec, s));
        } else {
            body(first);
        }
        read_sc();
        bool changed = false;
        any = false;
        for (int k = 0; k < nseg; ++k) {
            if (state[k] == kSegFinishing) { state[k] = kSegDone; changed = true; continue; }
            if (state[k] != kSegActive) continue;
            const double rr = hsc[16 * k + 4], bb = hsc[16 * k + 7];
            if (std::isnan(rr)) throw NanError("NaN encountered at iteration " + std::to_string(it));
            its[k] = it;
            if (it >= maxIter || !(rr > tol2 * bb)) { state[k] = kSegFinishing; changed = true; }
            else any = true;
        }
        if (changed) h2d(dstate.data(), state.data(), state.size(), s);
        first = false;
    }
    // the deferred x update of the segments that stopped at the last iteration
    if (std::any_of(state.begin(), state.end(), [](int v) { return v == kSegFinishing; }))
        launch_seg_finish_x(g, x.data(), d, sc.data(), dstate.data(), P, nseg, s);
    // compliance <f, u> per segment and the gradient, folded back onto the design in segment order
    launch_seg_dot(g, f.data(), x.data(), part.data(), P, nseg, sc.data(), 8, s);
    DBuf<double> &gst = B.gst;
    launch_compliance_gradient(g, S->K0.data(), S->rho.data(), x.data(), gst.data(), S->law, S->E0, S->Emin, S->gamma,
                               S->q, S->grav, S->h[0] * S->h[1] * S->h[2], 0, s);
    launch_fold_gradient(bs->g, g, P, nseg, B.dtop.data(), gst.data(), gsum, s);
    read_sc();
    for (int k = 0; k < nseg; ++k) comp[k] = hsc[16 * k + 8];
    sync(s);
}

int vf_lbl_run_zero_batched(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations,
                            int64_t smoothing, int fullMultigrid, int precision, int64_t pipelines, int useGraphs,
                            double *objective, double *grad, int gradWhere, double *perLayerCompliance,
                            int64_t *perLayerIterations) {
    VF_API_BEGIN
    need(mg && objective && grad, "null argument");
    need(layerIncrement >= 1, "layerIncrement must be positive");
    vf_sim *sim = mg->sim;
    const int L = mg->L;
    if (L < 1 || mgIterations < 1 || smoothing < 1 || !sim->walsh) {
        // the stacked form needs a multigrid preconditioner and the Walsh-basis element product: the
        // sequential evaluator gives the same results
        lbl_run(mg, layerIncrement, maxIter, tol, mgIterations, smoothing, fullMultigrid, precision, VF_LBL_INIT_ZERO,
                0, true, nullptr, false, VF_HOST, objective, grad, gradWhere, perLayerCompliance, perLayerIterations);
    } else {
        const int64_t numLayers = sim->ne[1];
        std::vector<int> heights;
        for (int64_t l = numLayers; l > 0; l -= std::min(layerIncrement, l)) heights.push_back((int)l);
        const int nl = (int)heights.size();
        hipStream_t s = mg->st();
        // the design's moduli without any mask
        const size_t ne = sim->nel();
        DBuf<double> Efull(ne);
        DBuf<float> E32full(ne);
        {
            Grid gf = sim->g;
            gf.ymask = gf.ne[1];
            launch_young(gf, sim->rho.data(), Efull.data(), E32full.data(), sim->law, sim->E0, sim->Emin, sim->gamma,
                         sim->q, 0, s);
        }
        // highest fine row at which any level carries a constraint: a batch grid is cut below the design's
        // top only above it
        int bcRow = -1;
        for (int l = 0; l <= L; ++l) {
            const std::vector<uint8_t> &hm = l == 0 ? sim->hmask : mg->lv[l]->hmask;
            const Grid &gl = mg->lv[l]->g;
            for (size_t n = 0; n < hm.size(); ++n)
                if (hm[n]) bcRow = std::max(bcRow, (int)((n / gl.nn[2]) % gl.nn[1]) << l);
        }
        const int G = 1 << L, P = (int)sim->ne[0] + G;
        // Batches: consecutive structures (tallest first), at most `pipelines` of them (0: no limit) and 2^25
        // stacked nodes (so that one FP64 vector stays below 1 GiB, the limit of the 32-bit buffer offsets of
        // the plane-marching kernels). The split minimises the sum over batches of (2^20 + stacked nodes): a
        // batch costs a fixed amount (setup, and per iteration the launch latency of ~100 kernels, measured
        // worth about 2^20 nodes of work) plus its grid, which is as tall as its tallest structure.
        const int64_t maxNodes = (int64_t)1 << 25, fixedNodes = (int64_t)1 << 20;
        auto rows_of = [&](int h) {
            int r = std::min<int>((int)sim->ne[1], (h + G - 1) / G * G);
            if (bcRow > r) r = (int)sim->ne[1];
            return r;
        };
        auto seg_nodes = [&](int rows) { return (int64_t)P * (rows + 1) * (sim->ne[2] + 1); };
        std::vector<int64_t> best(nl + 1, INT64_MAX);
        std::vector<int> cut(nl + 1, nl);
        best[nl] = 0;
        for (int i = nl - 1; i >= 0; --i) {
            const int64_t per = seg_nodes(rows_of(heights[i]));
            for (int j = i + 1; j <= nl; ++j) {
                const int cnt = j - i;
                if (cnt > 1 && (cnt * per > maxNodes || (pipelines > 0 && cnt > pipelines))) break;
                const int64_t c = fixedNodes + cnt * per + best[j];
                if (c < best[i]) { best[i] = c; cut[i] = j; }
            }
        }
        std::vector<double> comp(nl, 0.0);
        std::vector<int64_t> its(nl, 0);
        std::vector<std::shared_ptr<StackedBatch>> next;  // this evaluation's batches (the cache afterwards)
        DBuf<double> gsum(ne);
        launch_fill<double>(ne, gsum.data(), 0.0, s);
        for (int k0 = 0; k0 < nl;) {
            const int k1 = cut[k0];
            std::vector<int> top(heights.begin() + k0, heights.begin() + k1);
            lbl_stacked_batch(mg, next, top, rows_of(heights[k0]), maxIter, tol, (int)mgIterations, (int)smoothing,
                              fullMultigrid != 0, precision, useGraphs != 0, Efull.data(), E32full.data(),
                              comp.data() + k0, its.data() + k0, gsum.data());
            k0 = k1;
        }
        mg->stackCache = std::move(next);
        double total = 0;
        for (int k = 0; k < nl; ++k) {
            total += comp[k];
            if (perLayerCompliance) perLayerCompliance[k] = comp[k];
            if (perLayerIterations) perLayerIterations[k] = its[k];
        }
        *objective = 0.5 * total / (double)nl;
        std::vector<double> hg(ne);
        d2h(hg.data(), gsum.data(), ne, s);
        sync(s);
        for (auto &v : hg) v /= (double)nl;
        if (gradWhere == VF_DEVICE) { h2d(grad, hg.data(), ne, s); sync(s); }
        else std::copy(hg.begin(), hg.end(), grad);
        // the simulator ends, like the sequential evaluator, at the last (lowest) layer's mask
        sim->set_mask_height(sim->h[1] * (double)heights.back(), true);
    }
    VF_API_END
}

int vf_smoothing_filter(const int64_t dims[3], double rmin, int type, const double *in, double *out, int where,
                        void *stream) {
    VF_API_BEGIN
    need(dims && in && out, "null argument");
    need(rmin > 0, "filter radius must be positive");
    need(in != out, "smoothing filter cannot run in place");
    const int d[3] = {(int)dims[0], (int)dims[1], (int)dims[2]};
    const size_t n = (size_t)d[0] * d[1] * d[2];
    with_buffers(n, n, in, out, where, as_stream(stream), [&](const double *i, const double *, double *o) {
        launch_smoothing_filter(d, rmin, type, i, o, as_stream(stream));
    });
    VF_API_END
}

int vf_smoothing_filter_2d(const int64_t dims[2], double rmin, int type, const double *in, double *out, int where,
                           void *stream) {
    VF_API_BEGIN
    need(dims && in && out, "null argument");
    need(rmin > 0, "filter radius must be positive");
    need(in != out, "smoothing filter cannot run in place");
    need(dims[0] >= 1 && dims[1] >= 1 && dims[0] * dims[1] < INT_MAX, "invalid grid dimensions");
    const int d[2] = {(int)dims[0], (int)dims[1]};
    const size_t n = (size_t)d[0] * d[1];
    with_buffers(n, n, in, out, where, as_stream(stream), [&](const double *i, const double *, double *o) {
        vf::planar::launch_smoothing_filter_2d(d, rmin, type, i, o, as_stream(stream));
    });
    VF_API_END
}

static void check_grid_dims(int64_t nd, const int64_t *dims, int64_t minimum) {
    need(dims != nullptr, "null argument");
    need(nd == 2 || nd == 3, "grid dimension count must be 2 or 3");
    int64_t n = 1;
    for (int k = 0; k < nd; ++k) {
        need(dims[k] >= minimum, "grid too small for this filter");
        n *= dims[k];
    }
    need(n < INT_MAX, "grid too large");
}

static int upsample_impl(int64_t nd, const int64_t *inDims, int64_t factor, const double *a, double *b, bool adjoint,
                         int where, void *stream) {
    VF_API_BEGIN
    check_grid_dims(nd, inDims, 2);
    need(a && b && a != b, "null or aliased argument");
    need(factor >= 1 && factor < 4096, "upsampling factor must be positive");
    int64_t od[3];
    upsample_dims((int)nd, inDims, (int)factor, od);
    size_t nin = 1, nout = 1;
    for (int k = 0; k < nd; ++k) {
        nin *= (size_t)inDims[k];
        nout *= (size_t)od[k];
    }
    need(nout < (size_t)INT_MAX, "grid too large");
    with_buffers(adjoint ? nout : nin, adjoint ? nin : nout, a, b, where, as_stream(stream),
                 [&](const double *i, const double *, double *o) {
                     launch_upsample((int)nd, inDims, (int)factor, i, o, adjoint, as_stream(stream));
                 });
    VF_API_END
}
int vf_upsample_filter(int64_t nd, const int64_t *inDims, int64_t factor, const double *in, double *out, int where,
                       void *stream) {
    return upsample_impl(nd, inDims, factor, in, out, false, where, stream);
}
int vf_upsample_filter_backprop(int64_t nd, const int64_t *inDims, int64_t factor, const double *gOut, double *gIn,
                                int where, void *stream) {
    return upsample_impl(nd, inDims, factor, gOut, gIn, true, where, stream);
}

static int v2c_impl(int64_t nd, const int64_t *inDims, const double *a, double *b, bool adjoint, int where,
                    void *stream) {
    VF_API_BEGIN
    check_grid_dims(nd, inDims, 2);
    need(a && b && a != b, "null or aliased argument");
    size_t nin = 1, nout = 1;
    for (int k = 0; k < nd; ++k) {
        nin *= (size_t)inDims[k];
        nout *= (size_t)(inDims[k] - 1);
    }
    with_buffers(adjoint ? nout : nin, adjoint ? nin : nout, a, b, where, as_stream(stream),
                 [&](const double *i, const double *, double *o) {
                     launch_vertex_to_cell((int)nd, inDims, i, o, adjoint, as_stream(stream));
                 });
    VF_API_END
}
int vf_vertex_to_cell(int64_t nd, const int64_t *inDims, const double *in, double *out, int where, void *stream) {
    return v2c_impl(nd, inDims, in, out, false, where, stream);
}
int vf_vertex_to_cell_backprop(int64_t nd, const int64_t *inDims, const double *gOut, double *gIn, int where,
                               void *stream) {
    return v2c_impl(nd, inDims, gOut, gIn, true, where, stream);
}

int vf_langelaar(const int64_t dims[2], const double *in, double *out, double *smax, void *stream) {
    VF_API_BEGIN
    check_grid_dims(2, dims, 1);
    need(in && out && smax && in != out && smax != in && smax != out, "null or aliased argument");
    launch_langelaar((int)dims[0], (int)dims[1], in, out, smax, as_stream(stream));
    VF_API_END
}
int vf_langelaar_backprop(const int64_t dims[2], const double *g, const double *vars, const double *filtered,
                          const double *smax, double *scratch, double *out, void *stream) {
    VF_API_BEGIN
    check_grid_dims(2, dims, 1);
    need(g && vars && filtered && smax && scratch && out, "null argument");
    need(out != g && out != vars && out != filtered && out != smax && out != scratch && scratch != g &&
             scratch != vars && scratch != filtered && scratch != smax,
         "out and scratch must not alias the inputs or each other");
    launch_langelaar_backprop((int)dims[0], (int)dims[1], g, vars, filtered, smax, scratch, out, as_stream(stream));
    VF_API_END
}

int vf_projection(double beta, int64_t n, const double *in, double *out, int where, void *stream) {
    VF_API_BEGIN
    need(in && out, "null argument");
    with_buffers(n, n, in, out, where, as_stream(stream), [&](const double *i, const double *, double *o) {
        launch_projection(beta, n, i, o, as_stream(stream));
    });
    VF_API_END
}

int vf_projection_backprop(double beta, int64_t n, const double *g, const double *vars, double *out, int where,
                           void *stream) {
    VF_API_BEGIN
    need(g && vars && out, "null argument");
    with_buffers(
        n, n, g, out, where, as_stream(stream),
        [&](const double *i, const double *v, double *o) { launch_projection_backprop(beta, n, i, v, o, as_stream(stream)); },
        vars);
    VF_API_END
}

int vf_oc_update(int64_t n, const double *x0, const double *dJ, const double *dc, double lambda, double move, double p,
                 double *out, int where, void *stream) {
    VF_API_BEGIN
    need(x0 && dJ && dc && out, "null argument");
    hipStream_t s = as_stream(stream);
    if (where == VF_DEVICE) {
        launch_oc_update(n, x0, dJ, dc, lambda, move, p, out, s);
    } else {
        DBuf<double> a(n), b(n), c(n), o(n);
        h2d(a.data(), x0, n, s); h2d(b.data(), dJ, n, s); h2d(c.data(), dc, n, s);
        launch_oc_update(n, a.data(), b.data(), c.data(), lambda, move, p, o.data(), s);
        d2h(out, o.data(), n, s);
        sync(s);
    }
    VF_API_END
}

int vf_mean(int64_t n, const double *x, int where, void *stream, double *out) {
    VF_API_BEGIN
    need(x && out && n > 0, "invalid argument");
    hipStream_t s = as_stream(stream);
    DBuf<double> part(reduction_partials_full(n) + 1), res(1), tmp(where == VF_HOST ? n : 0);
    const double *dx = x;
    if (where == VF_HOST) { h2d(tmp.data(), x, n, s); dx = tmp.data(); }
    launch_dot_full(n, dx, nullptr, part.data(), res.data(), s);
    double v;
    d2h(&v, res.data(), 1, s);
    sync(s);
    *out = v / (double)n;
    VF_API_END
}

int vf_mean_noalloc(int64_t n, const double *x, void *stream, double *out, int outWhere) {
    VF_API_BEGIN
    need(x && out && n > 0, "invalid argument");
    if (n > NOALLOC_MAX_N)
        throw std::invalid_argument("vf_mean_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                                    std::to_string(NOALLOC_MAX_N) + " values its fixed scratch array covers");
    need(reduction_partials_full((size_t)n) <= NOALLOC_PARTIALS, "vf_mean_noalloc: partial-sum count out of range");
    hipStream_t s = as_stream(stream);
    double *sc = scratch_on_current_device();
    // every partial the final sum reads is written by the first pass: nothing needs clearing
    launch_dot_full((size_t)n, x, nullptr, sc, sc + NOALLOC_PARTIALS, s);
    if (outWhere == VF_DEVICE) {
        launch_div_scalar(sc + NOALLOC_PARTIALS, (double)n, out, s);
    } else {
        double v;
        d2h(&v, sc + NOALLOC_PARTIALS, 1, s);
        sync(s);
        *out = v / (double)n;
    }
    VF_API_END
}

// ---------------------------------------------------------------------------------------------- bench
int vf_bench_apply_K(vf_sim *sim, int64_t reps, int precision, double *avg_ms) {
    VF_API_BEGIN
    need(sim && avg_ms && reps > 0, "invalid argument");
    need(!sim->p2, "the kernel benchmarks cover the 3D kernels");
    sim->refresh_grid();
    const size_t n = 3 * sim->nn();
    hipStream_t s = sim->stream;
    hipEvent_t a, b;
    VF_HIP(hipEventCreate(&a));
    VF_HIP(hipEventCreate(&b));
    float ms = 0;
    if (precision == VF_PREC_FP64) {
        double *u = sim->stage_in(n), *y = sim->stage_out(n);
        launch_fill<double>(n, u, 1e-3, s);
        sim_apply_K<double>(sim, u, y, s);
        VF_HIP(hipEventRecord(a, s));
        for (int64_t i = 0; i < reps; ++i)
            sim_apply_K<double>(sim, u, y, s);
        VF_HIP(hipEventRecord(b, s));
    } else {
        sim->s32a.resize(n);
        sim->s32b.resize(n);
        float *u = sim->s32a.data(), *y = sim->s32b.data();
        launch_fill<float>(n, u, 1e-3f, s);
        sim_apply_K<float>(sim, u, y, s);
        VF_HIP(hipEventRecord(a, s));
        for (int64_t i = 0; i < reps; ++i)
            sim_apply_K<float>(sim, u, y, s);
        VF_HIP(hipEventRecord(b, s));
    }
    VF_HIP(hipEventSynchronize(b));
    VF_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = (double)ms / (double)reps;
    VF_API_END
}

int vf_mg_set_use_graphs(vf_mg *mg, int enable) {
    VF_API_BEGIN
    need(mg != nullptr, "null solver");
    mg->useGraphs = enable != 0;
    VF_API_END
}

int vf_mg_set_kernel_timing(vf_mg *mg, int enable) {
    VF_API_BEGIN
    need(mg != nullptr, "null solver");
    if (enable)
        for (hipEvent_t &e : mg->kdEv)
            if (!e) VF_HIP(hipEventCreate(&e));
    mg->timeKd = enable != 0;
    mg->kdMs = 0;
    mg->kdCount = 0;
    VF_API_END
}

int vf_mg_kernel_timing(vf_mg *mg, int64_t *count, double *total_ms) {
    VF_API_BEGIN
    need(mg && count && total_ms, "null argument");
    *count = mg->kdCount;
    *total_ms = mg->kdMs;
    VF_API_END
}

int vf_mg_tuning(vf_mg *mg, int precision, int choice[2], double ms[6]) {
    VF_API_BEGIN
    need(mg && choice && ms, "null argument");
    need(!mg->p2, "planar grids have no autotuned kernel families");
    need(precision == VF_PREC_FP64 || precision == VF_PREC_FP32, "precision must be VF_PREC_FP64 or VF_PREC_FP32");
    mg->sync_masks();
    mg->ensure_stiffness();
    const int p = precision == VF_PREC_FP64 ? 0 : 1;
    if (p == 0) {
        mg->tune<double>();
        choice[0] = mg->ku_family<double>();
        choice[1] = mg->gs_family<double>();
    } else {
        mg->tune<float>();
        choice[0] = mg->ku_family<float>();
        choice[1] = mg->gs_family<float>();
    }
    for (int i = 0; i < 6; ++i) ms[i] = mg->tuneMs[p][i];
    VF_API_END
}

int vf_mg_tuning_level1(vf_mg *mg, int precision, int choice[2], double ms[6]) {
    VF_API_BEGIN
    need(mg && choice && ms, "null argument");
    need(!mg->p2, "planar grids have no autotuned kernel families");
    need(precision == VF_PREC_FP64 || precision == VF_PREC_FP32, "precision must be VF_PREC_FP64 or VF_PREC_FP32");
    mg->sync_masks();
    mg->ensure_stiffness();
    const int p = precision == VF_PREC_FP64 ? 0 : 1;
    if (p == 0) mg->tune<double>();
    else mg->tune<float>();
    choice[0] = mg->l1_form(p, 0);
    choice[1] = mg->l1_form(p, 1);
    for (int i = 0; i < 6; ++i) ms[i] = mg->tuneMs[p][6 + i];
    VF_API_END
}

int vf_bench_mg_kernel(vf_mg *mg, int which, int64_t level, int64_t reps, int precision, double *avg_ms) {
    VF_API_BEGIN
    need(mg && avg_ms && reps > 0, "invalid argument");
    need(!mg->p2, "the kernel benchmarks cover the 3D kernels");
    need(level >= 0 && level < mg->L, "level must be a smoothed level (0 <= l < numLevels)");
    need(which >= 0 && which <= 2, "which must be 0 (K u), 1 (residual) or 2 (one smoothing sweep)");
    need(precision == VF_PREC_FP64 || precision == VF_PREC_FP32, "precision must be VF_PREC_FP64 or VF_PREC_FP32");
    mg->sync_masks();
    mg->ensure_stiffness();
    const int l = (int)level;
    const size_t n = 3 * (size_t)mg->lv[l]->g.nnodes;
    hipStream_t s = mg->st();
    hipEvent_t a, b;
    VF_HIP(hipEventCreate(&a));
    VF_HIP(hipEventCreate(&b));
    auto run = [&](auto zero) {
        using T = decltype(zero);
        mg->template tune<T>();
        DBuf<T> u(n), y(n), rhs(n);
        launch_fill<T>(n, u.data(), T(1e-3), s);
        launch_fill<T>(n, rhs.data(), T(1), s);
        auto once = [&]() {
            if (which == 0) mg->apply<T>(l, u.data(), y.data(), nullptr, AM_SET, false);
            else if (which == 1) mg->apply<T>(l, u.data(), y.data(), rhs.data(), AM_RESIDUAL, true);
            else if (mg->template paired<T>(l)) mg->sweep_to<T>(l, u.data(), y.data(), rhs.data(), true);
            else mg->smooth<T>(l, u.data(), rhs.data(), true);
        };
        once();
        VF_HIP(hipEventRecord(a, s));
        for (int64_t i = 0; i < reps; ++i) once();
        VF_HIP(hipEventRecord(b, s));
        VF_HIP(hipEventSynchronize(b));
    };
    if (precision == VF_PREC_FP64) run(0.0);
    else run(0.0f);
    float ms = 0;
    VF_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = (double)ms / (double)reps;
    VF_API_END
}

}  // extern "C"
