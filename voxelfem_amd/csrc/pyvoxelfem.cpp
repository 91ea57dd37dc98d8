// pyvoxelfem.cpp -- pybind11 module `pyVoxelFEM` with the reference's Python API
// (python_bindings/VoxelFEM.cc), implemented in C++ on top of the C ABI of include/voxelfem.h.
//
// The [1,1,1] (trilinear hex) and [1,1] (bilinear quad) double-precision simulators are provided, as in
// the reference. Fields cross this boundary as numpy arrays exactly like the reference: nodal fields are
// (numNodes, N) float64 arrays (N = 3 or 2), element fields flat C-order float64 vectors. Device-resident state
// (densities, Young's moduli, displacements, load vectors, filter intermediates) stays in HBM between
// calls; host copies are made only when a value is handed back to Python.
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cmath>
#include <cstring>
#include <chrono>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "voxelfem.h"

namespace py = pybind11;

namespace {

void check(int rc) {
    if (rc != VF_OK) {
        std::string msg = vf_last_error();
        if (rc == VF_ERR_NAN) throw std::logic_error(msg);
        throw std::runtime_error(msg);
    }
}

using Arr = py::array_t<double, py::array::forcecast>;

// ---- device vector (RAII over vf_malloc) ---------------------------------------------------------
struct DevVec {
    double *p = nullptr;
    size_t n = 0;
    DevVec() = default;
    explicit DevVec(size_t count) { resize(count); }
    DevVec(const DevVec &) = delete;
    DevVec &operator=(const DevVec &) = delete;
    ~DevVec() { if (p) vf_free(p); }
    void resize(size_t count) {
        if (count == n) return;
        if (p) vf_free(p);
        p = nullptr;
        n = 0;
        if (count) check(vf_malloc((void **)&p, count * sizeof(double)));
        n = count;
    }
    void upload(const double *h, size_t count) {
        resize(count);
        check(vf_memcpy(p, h, count * sizeof(double), 0, nullptr));
    }
    void download(double *h) const { check(vf_memcpy(h, p, n * sizeof(double), 1, nullptr)); }
    void copy_from(const DevVec &o) {
        resize(o.n);
        check(vf_memcpy(p, o.p, n * sizeof(double), 2, nullptr));
    }
    std::vector<double> host() const {
        std::vector<double> v(n);
        if (n) download(v.data());
        return v;
    }
    void fill_host(double v) {
        std::vector<double> h(n, v);
        if (n) check(vf_memcpy(p, h.data(), n * sizeof(double), 0, nullptr));
    }
};

// nodal (nn, N) numpy <-> SoA component planes (N = 3, or 2 on a planar grid)
std::vector<double> vfield_to_soa(const Arr &a, size_t nn, int N = 3) {
    if (a.ndim() != 2 || (size_t)a.shape(0) != nn || a.shape(1) != N)
        throw std::runtime_error("expected a nodal field of shape (" + std::to_string(nn) + ", " + std::to_string(N) + ")");
    auto r = a.unchecked<2>();
    std::vector<double> out((size_t)N * nn);
    for (size_t n = 0; n < nn; ++n)
        for (int c = 0; c < N; ++c) out[c * nn + n] = r(n, c);
    return out;
}
py::array_t<double> soa_to_vfield(const double *p, size_t nn, int N = 3) {
    py::array_t<double, py::array::f_style> out({(py::ssize_t)nn, (py::ssize_t)N});
    std::memcpy(out.mutable_data(), p, (size_t)N * nn * sizeof(double));
    return out;
}
std::vector<double> to_vec(const Arr &a, size_t n, const char *what) {
    if ((size_t)a.size() != n) throw std::runtime_error(std::string("Incorrect ") + what + " size");
    std::vector<double> v(n);
    std::memcpy(v.data(), a.data(), n * sizeof(double));
    return v;
}
py::array_t<double> from_vec(const std::vector<double> &v) {
    py::array_t<double> out((py::ssize_t)v.size());
    std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(double));
    return out;
}

enum class InterpolationLaw { SIMP = 0, RAMP = 1 };
enum class NumberType { DOUBLE = 0, FLOAT = 1 };

// ====================================================================================================
// TensorProductSimulator<double,1,1,1>
// ====================================================================================================
// A planar (2D) simulator is held with N = 2 and a third axis of one element (ne[2] = 1, no node along it:
// nnodes3()[2] = 1), so node and element indices read the same as the 3D ones with iz = ez = 0.
struct TPS : std::enable_shared_from_this<TPS> {
    vf_sim *sim = nullptr;
    int N = 3;
    std::array<int64_t, 3> ne;
    std::array<double, 3> lo, hi;
    size_t nn = 0, nel = 0;
    std::vector<uint8_t> dmask;
    std::vector<double> dvals, forces;
    double young = 1, poisson = 0;
    InterpolationLaw law = InterpolationLaw::SIMP;
    double E0 = 1, Emin = 1e-4, gamma = 3, q = 3;
    std::array<double, 3> grav{0, 0, 0};
    std::shared_ptr<struct MG> directSolver;  // 0-level hierarchy backing solve(f)

    void *stream = nullptr;  // the simulator's own stream (ordered with the null stream)

    TPS(const std::array<double, 3> &minC, const std::array<double, 3> &maxC, const std::array<int64_t, 3> &n, int dim = 3)
        : N(dim), ne(n), lo(minC), hi(maxC) {
        check(vf_stream_create(&stream));
        if (N == 2) {
            ne[2] = 1;
            lo[2] = 0;
            hi[2] = 1;
            check(vf_sim_create_2d(&sim, ne.data(), lo.data(), hi.data(), stream));
        } else {
            check(vf_sim_create(&sim, ne.data(), lo.data(), hi.data(), stream));
        }
        const auto nd = nnodes3();
        nn = (size_t)(nd[0] * nd[1] * nd[2]);
        nel = (size_t)ne[0] * ne[1] * ne[2];
        dmask.assign(nn, 0);
        dvals.assign((size_t)N * nn, 0.0);
        forces.assign((size_t)N * nn, 0.0);
        push_law();
    }
    uint8_t fullMask() const { return N == 2 ? 3 : 7; }
    // the first N entries of a 3-array, for Python
    template <typename T>
    std::vector<T> dims(const std::array<T, 3> &a) const { return std::vector<T>(a.begin(), a.begin() + N); }
    ~TPS() {
        if (sim) vf_sim_destroy(sim);
        if (stream) vf_stream_destroy(stream);
    }

    std::array<int64_t, 3> nnodes3() const { return {ne[0] + 1, ne[1] + 1, N == 2 ? 1 : ne[2] + 1}; }
    std::array<double, 3> h() const { return {(hi[0] - lo[0]) / ne[0], (hi[1] - lo[1]) / ne[1], (hi[2] - lo[2]) / ne[2]}; }
    // element stretchings (the node spacing h() unless set through the dx property)
    std::array<double, 3> stretch() const {
        std::array<double, 3> s;
        check(vf_sim_get_stretchings(sim, s.data()));
        return s;
    }
    std::array<double, 3> nodePos(size_t n) const {
        const auto N = nnodes3();
        const size_t iz = n % N[2], iy = (n / N[2]) % N[1], ix = n / (N[1] * N[2]);
        const auto hh = h();
        return {lo[0] + ix * hh[0], lo[1] + iy * hh[1], lo[2] + iz * hh[2]};
    }
    void push_law() { check(vf_sim_set_law(sim, (int)law, E0, Emin, gamma, q)); }
    void push_bc() {
        check(vf_sim_set_dirichlet(sim, dmask.data(), dvals.data()));
        check(vf_sim_set_forces(sim, forces.data()));
    }
    bool hasBC() const {
        for (auto m : dmask) if (m) return true;
        for (auto f : forces) if (f != 0) return true;
        return false;
    }
    // BCBuilder::setDirichlet (TensorProductSimulator.hh:488-504)
    void setDirichlet(size_t n, const std::array<double, 3> &val, uint8_t cm) {
        for (int c = 0; c < N; ++c) {
            if (!((cm >> c) & 1)) continue;
            if (!((dmask[n] >> c) & 1)) {
                dmask[n] |= (uint8_t)(1 << c);
                dvals[c * nn + n] = val[c];
            } else if (std::abs(dvals[c * nn + n] - val[c]) > 1e-10) {
                throw std::runtime_error("Conflicting dirichlet displacements.");
            }
        }
    }

    std::vector<double> getDensities() const {
        std::vector<double> v(nel);
        check(vf_sim_get_densities(sim, v.data(), VF_HOST));
        return v;
    }
    void setDensities(const std::vector<double> &v) { check(vf_sim_set_densities(sim, v.data(), VF_HOST)); }
    std::vector<double> getE() const {
        std::vector<double> v(nel);
        check(vf_sim_get_young(sim, v.data(), VF_HOST));
        return v;
    }
    int keSize() const { return N == 2 ? 8 : 24; }
    std::vector<double> K0() const {
        std::vector<double> k(576);
        check(vf_sim_get_K0(sim, k.data()));
        return k;
    }
    double maskHeight() const {
        double hgt;
        check(vf_sim_get_mask_info(sim, nullptr, nullptr, &hgt));
        return hgt;
    }
    int64_t firstMaskedLayer() const {
        int64_t m;
        check(vf_sim_get_mask_info(sim, &m, nullptr, nullptr));
        return m;
    }
    // local node l (2D: l = 2 i + j, mapped onto the 3D numbering 4 i + 2 j)
    size_t elemNode(size_t e, int l) const {
        if (N == 2) l *= 2;
        const size_t ez = e % ne[2], ey = (e / ne[2]) % ne[1], ex = e / (ne[1] * ne[2]);
        const auto N = nnodes3();
        return ((ex + ((l >> 2) & 1)) * N[1] + ey + ((l >> 1) & 1)) * N[2] + ez + (l & 1);
    }
    std::vector<double> solve(const std::vector<double> &fsoa);
};
using TPSPtr = std::shared_ptr<TPS>;

// ====================================================================================================
// MultigridSolver<double,1,1,1>
// ====================================================================================================
struct MG {
    TPSPtr tps;
    vf_mg *mg = nullptr;
    int64_t levels = 0;
    bool useGraphs = false;
    bool partialUpdates = true;
    MG(TPSPtr t, int64_t L) : tps(std::move(t)), levels(L) { check(vf_mg_create(&mg, tps->sim, L)); }
    ~MG() { if (mg) vf_mg_destroy(mg); }
    size_t nodes(int64_t l) const {
        if (l < 0 || l > levels) throw std::out_of_range("level out of range");
        int64_t n3[3];
        check(vf_mg_level_nodes(mg, l, n3));
        return (size_t)(n3[0] * n3[1] * n3[2]);
    }
};
using MGPtr = std::shared_ptr<MG>;

std::vector<double> TPS::solve(const std::vector<double> &fsoa) {
    if (!directSolver) directSolver = std::make_shared<MG>(shared_from_this(), 0);
    std::vector<double> x((size_t)N * nn, 0.0);
    check(vf_mg_update_stiffness(directSolver->mg));
    check(vf_mg_coarse_solve(directSolver->mg, fsoa.data(), x.data(), VF_HOST));
    return x;
}

// A coarse level of a hierarchy, returned by MultigridSolver.getSimulator(l > 0)
struct LevelView {
    MGPtr owner;
    int64_t l;
};

// ====================================================================================================
// Filters (TopologyOptimizationFilter.hh:90-400) -- device-resident
// ====================================================================================================
// A filter maps a grid of inDims values to one of outDims values (TopologyOptimizationFilter.hh:16-88); most keep
// the grid. dims: 2 or 3 entries (planar or 3D grid), C order.
struct Filter {
    std::vector<int64_t> inDims, outDims;
    bool dimsSet = false;
    virtual ~Filter() = default;
    virtual std::string name() const = 0;
    static size_t prod(const std::vector<int64_t> &d) {
        size_t p = 1;
        for (auto v : d) p *= (size_t)v;
        return p;
    }
    static void checkCount(const std::vector<int64_t> &d) {
        if (d.size() != 3 && d.size() != 2) throw std::runtime_error("Dimension count mismatch");
    }
    size_t inSize() const { return prod(inDims); }
    size_t outSize() const { return prod(outDims); }
    size_t size() const { return outSize(); }
    int nd() const { return (int)outDims.size(); }
    // the grid as three extents (a planar grid: {nx, ny, 1})
    static std::array<int64_t, 3> pad3(const std::vector<int64_t> &d) { return {d[0], d[1], d.size() == 3 ? d[2] : 1}; }
    void setInputDims(const std::vector<int64_t> &d) {
        checkCount(d);
        setInput(d);
        dimsSet = true;
    }
    void setOutputDims(const std::vector<int64_t> &d) {
        checkCount(d);
        setOutput(d);
        dimsSet = true;
    }
    void checkDims() const {
        if (!dimsSet)
            throw std::runtime_error(
                "Filter grid dimensions not set. Initialize a TopologyOptimizationProblem object with this filter "
                "before using it.");
    }
    virtual void apply(const DevVec &in, DevVec &out) = 0;
    virtual void backprop(const DevVec &g, const DevVec &vars, DevVec &out) = 0;

protected:
    virtual void setInput(const std::vector<int64_t> &d) { inDims = outDims = d; }
    virtual void setOutput(const std::vector<int64_t> &d) { inDims = outDims = d; }
};
using FilterPtr = std::shared_ptr<Filter>;

struct SmoothingFilter : Filter {
    enum class Type { Const = 0, Linear = 1 };
    int radius = 1;
    Type type = Type::Const;
    double rmin = 0;  // extension: a real kernel radius; 0 means radius + 1 (reference convention)
    SmoothingFilter(int r, Type t) : radius(r), type(t) {}
    std::string name() const override { return "SmoothingFilter"; }
    double kernelRadius() const { return rmin > 0 ? rmin : radius + 1.0; }
    void apply(const DevVec &in, DevVec &out) override {
        checkDims();
        out.resize(size());
        const auto d = pad3(outDims);
        if (nd() == 2) check(vf_smoothing_filter_2d(d.data(), kernelRadius(), (int)type, in.p, out.p, VF_DEVICE, nullptr));
        else check(vf_smoothing_filter(d.data(), kernelRadius(), (int)type, in.p, out.p, VF_DEVICE, nullptr));
    }
    void backprop(const DevVec &g, const DevVec &, DevVec &out) override { apply(g, out); }  // S^T = S
};

struct ProjectionFilter : Filter {
    double beta = 1.0;
    ProjectionFilter() = default;
    explicit ProjectionFilter(double b) : beta(b) {}
    std::string name() const override { return "ProjectionFilter"; }
    void apply(const DevVec &in, DevVec &out) override {
        checkDims();
        out.resize(in.n);
        check(vf_projection(beta, (int64_t)in.n, in.p, out.p, VF_DEVICE, nullptr));
    }
    void backprop(const DevVec &g, const DevVec &vars, DevVec &out) override {
        out.resize(g.n);
        check(vf_projection_backprop(beta, (int64_t)g.n, g.p, vars.p, out.p, VF_DEVICE, nullptr));
    }
    double invert(double v) const {
        if (v > 1.0 || v < 0.0)
            throw std::runtime_error("ProjectionFilter::invert domain error: target density for inversion is outside [0, 1].");
        return std::atanh((2 * v - 1) * std::tanh(0.5 * beta)) / beta + 0.5;
    }
};

struct PythonFilter : Filter {
    std::function<void(py::array_t<double>, py::array_t<double>)> apply_cb;
    std::function<void(py::array_t<double>, py::array_t<double>, py::array_t<double>)> backprop_cb;
    std::string name() const override { return "PythonFilter"; }
    void apply(const DevVec &in, DevVec &out) override {
        if (!apply_cb) throw std::runtime_error("Apply callback must be configured");
        auto hin = from_vec(in.host());
        py::array_t<double> res((py::ssize_t)outSize());
        apply_cb(hin, res);
        out.upload(res.data(), outSize());
    }
    void backprop(const DevVec &g, const DevVec &vars, DevVec &out) override {
        if (!backprop_cb) throw std::runtime_error("Backprop callback must be configured");
        auto hg = from_vec(g.host()), hv = from_vec(vars.host());
        py::array_t<double> res((py::ssize_t)inSize());
        backprop_cb(hg, hv, res);
        out.upload(res.data(), inSize());
    }
};

// UpsampleFilter (TopologyOptimizationFilter.hh:419-525): degree-1 interpolation onto a grid `factor` times finer
struct UpsampleFilter : Filter {
    int64_t factor;
    explicit UpsampleFilter(int64_t f) : factor(f) {
        if (f < 1) throw std::runtime_error("UpsampleFilter factor must be positive");
    }
    std::string name() const override { return "UpsampleFilter"; }
    void apply(const DevVec &in, DevVec &out) override {
        checkDims();
        out.resize(outSize());
        check(vf_upsample_filter((int64_t)nd(), inDims.data(), factor, in.p, out.p, VF_DEVICE, nullptr));
    }
    void backprop(const DevVec &g, const DevVec &, DevVec &out) override {
        checkDims();
        out.resize(inSize());
        check(vf_upsample_filter_backprop((int64_t)nd(), inDims.data(), factor, g.p, out.p, VF_DEVICE, nullptr));
    }

protected:
    void setInput(const std::vector<int64_t> &d) override {
        for (auto v : d)
            if (v < 2) throw std::runtime_error("Interpolation can only be applied to a 2^d grid or larger");
        inDims = d;
        outDims.resize(d.size());
        for (size_t k = 0; k < d.size(); ++k) outDims[k] = (d[k] - 1) * factor + 1;
    }
    void setOutput(const std::vector<int64_t> &d) override {
        for (auto v : d)
            if (v < 2) throw std::runtime_error("Interpolation can only be applied to a 2^d grid or larger");
        outDims = d;
        inDims.resize(d.size());
        for (size_t k = 0; k < d.size(); ++k) {
            inDims[k] = (d[k] - 1) / factor + 1;
            if ((inDims[k] - 1) * factor + 1 != d[k]) throw std::runtime_error("Output size is not divisible by factor");
        }
    }
};

// VertexToCellFilter (TopologyOptimizationFilter.hh:527-597): vertex values to cell means
struct VertexToCellFilter : Filter {
    std::string name() const override { return "VertexToCellFilter"; }
    void apply(const DevVec &in, DevVec &out) override {
        checkDims();
        out.resize(outSize());
        check(vf_vertex_to_cell((int64_t)nd(), inDims.data(), in.p, out.p, VF_DEVICE, nullptr));
    }
    void backprop(const DevVec &g, const DevVec &, DevVec &out) override {
        checkDims();
        out.resize(inSize());
        check(vf_vertex_to_cell_backprop((int64_t)nd(), inDims.data(), g.p, out.p, VF_DEVICE, nullptr));
    }

protected:
    void setInput(const std::vector<int64_t> &d) override {
        for (auto v : d)
            if (v < 2) throw std::runtime_error("Input grid must be 2^d or larger.");
        inDims = d;
        outDims.resize(d.size());
        for (size_t k = 0; k < d.size(); ++k) outDims[k] = d[k] - 1;
    }
    void setOutput(const std::vector<int64_t> &d) override {
        outDims = d;
        inDims.resize(d.size());
        for (size_t k = 0; k < d.size(); ++k) inDims[k] = d[k] + 1;
    }
};

// LangelaarFilter (TopologyOptimizationFilter.hh:600-712), build direction y. The filtered densities and the
// smooth maxima of the last apply are kept for backprop, as in the reference. Planar grids only: on 3D grids the
// reference's supporting region (NDVector.hh:212-229) holds the voxel itself, whose output it reads before
// computing it, so its result depends on the buffer's earlier contents.
struct LangelaarFilter : Filter {
    DevVec filtered, smax, scratch;
    std::string name() const override { return "LangelaarFilter"; }
    void apply(const DevVec &in, DevVec &out) override {
        checkDims();
        out.resize(outSize());
        smax.resize(outSize());
        check(vf_langelaar(outDims.data(), in.p, out.p, smax.p, nullptr));
        filtered.copy_from(out);
    }
    void backprop(const DevVec &g, const DevVec &vars, DevVec &out) override {
        checkDims();
        if (filtered.n != outSize()) throw std::runtime_error("LangelaarFilter: apply must run before backprop");
        out.resize(inSize());
        scratch.resize(inSize());
        check(vf_langelaar_backprop(outDims.data(), g.p, vars.p, filtered.p, smax.p, scratch.p, out.p, nullptr));
    }

protected:
    void setInput(const std::vector<int64_t> &d) override { setOutput(d); }
    void setOutput(const std::vector<int64_t> &d) override {
        if (d.size() != 2)
            throw std::runtime_error("LangelaarFilter is provided for planar grids: on 3D grids the reference's supporting "
                                     "region reads the voxel's own output before computing it");
        inDims = outDims = d;
    }
};

// FilterChain (TopologyOptimizationFilter.hh:90-187): the chain is sized from its output grid backwards, so
// grid-changing filters give a design grid different from the physical one
struct FilterChain {
    std::vector<FilterPtr> filters;
    std::vector<std::unique_ptr<DevVec>> vars;  // vars[0] design, vars.back() physical
    std::vector<int64_t> inDims, outDims;
    FilterChain(std::vector<FilterPtr> f, const std::vector<int64_t> &outGridDims) : filters(std::move(f)) {
        setOutputDimensions(outGridDims);
    }
    void allocate() {
        vars.clear();
        vars.emplace_back(std::make_unique<DevVec>(Filter::prod(inDims)));
        for (auto &fl : filters) vars.emplace_back(std::make_unique<DevVec>(fl->outSize()));
    }
    void setOutputDimensions(std::vector<int64_t> dims) {
        Filter::checkCount(dims);
        outDims = dims;
        for (size_t i = filters.size(); i-- > 0;) {
            filters[i]->setOutputDims(dims);
            dims = filters[i]->inDims;
        }
        inDims = dims;
        allocate();
    }
    void setInputDimensions(std::vector<int64_t> dims) {
        Filter::checkCount(dims);
        inDims = dims;
        for (auto &fl : filters) {
            fl->setInputDims(dims);
            dims = fl->outDims;
        }
        outDims = dims;
        allocate();
    }
    size_t size() const { return Filter::prod(inDims); }
    size_t numPhysicalVars() const { return Filter::prod(outDims); }
    DevVec &design() { return *vars.front(); }
    DevVec &physical() { return *vars.back(); }
    // setDesignVars (TopologyOptimizationFilter.hh:142-152)
    void setDesignVarsDevice(const DevVec &x) {
        if (x.n != size()) throw std::runtime_error("Variable size mismatch");
        design().copy_from(x);
        for (size_t i = 0; i < filters.size(); ++i) filters[i]->apply(*vars[i], *vars[i + 1]);
    }
    // applyInPlace (154-160)
    void applyInPlace(DevVec &x, DevVec &scratch) {
        for (auto &f : filters) {
            f->apply(x, scratch);
            std::swap(x.p, scratch.p);
            std::swap(x.n, scratch.n);
        }
    }
    // backprop (162-175)
    void backprop(DevVec &g) {
        DevVec scratch(g.n);
        for (size_t i = filters.size(); i-- > 0;) {
            filters[i]->backprop(g, *vars[i], scratch);
            std::swap(g.p, scratch.p);
            std::swap(g.n, scratch.n);
        }
    }
};
using FilterChainPtr = std::shared_ptr<FilterChain>;

// ====================================================================================================
// Constraints (TopologyOptimizationConstraint.hh:24-40)
// ====================================================================================================
struct Constraint {
    virtual ~Constraint() = default;
    virtual double evaluate(const DevVec &x) = 0;
    virtual void backprop(const DevVec &x, DevVec &out) = 0;
};
struct TotalVolumeConstraint : Constraint {
    double volumeFraction;
    explicit TotalVolumeConstraint(double v) : volumeFraction(v) {}
    double evaluate(const DevVec &x) override {
        double m;
        check(vf_mean((int64_t)x.n, x.p, VF_DEVICE, nullptr, &m));
        return 1.0 - m / volumeFraction;
    }
    void backprop(const DevVec &x, DevVec &out) override {
        out.resize(x.n);
        out.fill_host(-1.0 / (volumeFraction * (double)x.n));
    }
};
using ConstraintPtr = std::shared_ptr<Constraint>;

// ====================================================================================================
// Objectives (TopologyOptimizationObjective.hh:25-105)
// ====================================================================================================
struct Objective {
    virtual ~Objective() = default;
    virtual double evaluate() = 0;
    virtual void gradient(DevVec &out) = 0;
    virtual void updateCache(const DevVec &xPhys) = 0;
};
using ObjectivePtr = std::shared_ptr<Objective>;

struct ComplianceObjective : Objective {
    TPSPtr tps;
    DevVec u, f;
    explicit ComplianceObjective(TPSPtr t, bool skipSolve = false) : tps(std::move(t)) {
        f.resize((size_t)tps->N * tps->nn);
        u.resize((size_t)tps->N * tps->nn);
        u.fill_host(0.0);
        check(vf_sim_build_load(tps->sim, f.p, VF_DEVICE));
        check(vf_stream_sync(nullptr));
        if (!skipSolve) {
            DevVec rho(tps->nel);
            check(vf_sim_get_densities(tps->sim, rho.p, VF_DEVICE));
            updateCache(rho);
        }
    }
    double compliance() {
        double c;
        check(vf_sim_dot(tps->sim, f.p, u.p, 0, VF_DEVICE, &c));
        return 0.5 * c;
    }
    double evaluate() override { return compliance(); }
    void gradient(DevVec &out) override {
        out.resize(tps->nel);
        check(vf_sim_compliance_gradient(tps->sim, u.p, out.p, VF_DEVICE));
    }
    void updateCache(const DevVec &xPhys) override {
        check(vf_sim_set_densities(tps->sim, xPhys.p, VF_DEVICE));
        auto fh = f.host();
        auto x = tps->solve(fh);
        u.upload(x.data(), x.size());
    }
};

struct MultigridComplianceObjective : ComplianceObjective {
    MGPtr mg;
    int64_t cgIter = 100;
    double tol = 1e-5;
    int64_t mgIterations = 1, mgSmoothingIterations = 2;
    bool fullMultigrid = true;
    bool zeroInit = false;
    std::function<void(size_t, double)> residual_cb;
    int precision = VF_PREC_FP64;  // extension: FP32 V-cycle inside the FP64 PCG
    int64_t lastIterations = 0;
    explicit MultigridComplianceObjective(MGPtr m) : ComplianceObjective(m->tps, true), mg(std::move(m)) {
        DevVec rho(tps->nel);
        check(vf_sim_get_densities(tps->sim, rho.p, VF_DEVICE));
        updateCache(rho);
    }
    static int trampoline(void *user, int64_t it, const double *, const double *r) {
        auto *self = (MultigridComplianceObjective *)user;
        double s = 0;
        for (size_t i = 0; i < (size_t)self->tps->N * self->tps->nn; ++i) s += r[i] * r[i];
        py::gil_scoped_acquire g;
        self->residual_cb((size_t)it, std::sqrt(s));
        return 0;
    }
    void updateCache(const DevVec &xPhys) override {
        check(vf_sim_set_densities(tps->sim, xPhys.p, VF_DEVICE));
        if (zeroInit) u.fill_host(0.0);
        vf_pcg_stats st{};
        check(vf_mg_pcg(mg->mg, u.p, f.p, cgIter, tol, mgIterations, mgSmoothingIterations, fullMultigrid ? 1 : 0,
                        precision, 0, residual_cb ? &trampoline : nullptr, this, VF_DEVICE, &st));
        lastIterations = st.iterations;
    }
};

// ====================================================================================================
// TopologyOptimizationProblem (TopologyOptimizationProblem.hh:21-151)
// ====================================================================================================
struct Problem {
    TPSPtr tps;
    ObjectivePtr objective;
    std::vector<ConstraintPtr> constraints;
    std::vector<FilterPtr> filterList;
    FilterChainPtr chain;
    bool varsAreSet = false;

    Problem(TPSPtr t, ObjectivePtr o, std::vector<ConstraintPtr> c, std::vector<FilterPtr> f)
        : tps(std::move(t)), objective(std::move(o)), constraints(std::move(c)), filterList(f) {
        chain = std::make_shared<FilterChain>(f, std::vector<int64_t>(tps->ne.begin(), tps->ne.begin() + tps->N));
    }
    virtual ~Problem() = default;
    size_t numVars() const { return chain->size(); }
    void setVarsDevice(const DevVec &x) {
        chain->setDesignVarsDevice(x);
        objective->updateCache(chain->physical());
        varsAreSet = true;
    }
    virtual bool setVars(const std::vector<double> &x, bool) {
        DevVec d;
        d.upload(x.data(), x.size());
        setVarsDevice(d);
        return true;
    }
    // entry used by OCOptimizer: dispatches to a Python override when there is one
    virtual void setVarsFrom(const DevVec &x) { setVarsDevice(x); }
    virtual double evaluateObjective() {
        if (!varsAreSet) throw std::runtime_error("Must call setVars first!");
        return objective->evaluate();
    }
    void evaluateObjectiveGradientDevice(DevVec &g) {
        if (!varsAreSet) throw std::runtime_error("Must call setVars first!");
        objective->gradient(g);
        chain->backprop(g);
    }
    virtual std::vector<double> evaluateObjectiveGradientAndReturn() {
        DevVec g;
        evaluateObjectiveGradientDevice(g);
        return g.host();
    }
    std::vector<double> evaluateConstraints() {
        if (!varsAreSet) throw std::runtime_error("Must call setVars first!");
        std::vector<double> out;
        for (auto &c : constraints) out.push_back(c->evaluate(chain->physical()));
        return out;
    }
    void evaluateConstraintsJacobianDevice(std::vector<std::unique_ptr<DevVec>> &rows) {
        if (!varsAreSet) throw std::runtime_error("Must call setVars first!");
        rows.clear();
        for (auto &c : constraints) {
            auto g = std::make_unique<DevVec>();
            c->backprop(chain->physical(), *g);
            chain->backprop(*g);
            rows.push_back(std::move(g));
        }
    }
    double evaluateOCConstraintAtVars(DevVec &x, DevVec &scratch) {
        if (constraints.size() != 1 || !std::dynamic_pointer_cast<TotalVolumeConstraint>(constraints[0]))
            throw std::runtime_error("Applicable only for a topology optimization with a single (volume) constraint");
        chain->applyInPlace(x, scratch);
        return constraints[0]->evaluate(x);
    }
};

struct PyProblem : Problem {
    using Problem::Problem;
    bool setVars(const std::vector<double> &x, bool force) override {
        PYBIND11_OVERRIDE(bool, Problem, setVars, x, force);
    }
    double evaluateObjective() override { PYBIND11_OVERRIDE(double, Problem, evaluateObjective, ); }
    std::vector<double> evaluateObjectiveGradientAndReturn() override {
        PYBIND11_OVERRIDE_NAME(std::vector<double>, Problem, "evaluateObjectiveGradient", evaluateObjectiveGradientAndReturn, );
    }
    void setVarsFrom(const DevVec &x) override {
        py::gil_scoped_acquire gil;
        py::function ov = py::get_override(static_cast<const Problem *>(this), "setVars");
        if (ov) {
            auto h = x.host();
            ov(from_vec(h));
            return;
        }
        setVarsDevice(x);
    }
};

// ====================================================================================================
// OCOptimizer (OptimalityCriterion.hh:40-151)
// ====================================================================================================
struct OCOptimizer {
    Problem &p;
    double lambda_min = 1, lambda_max = 2;
    size_t loopCountGuard = 100;
    int64_t lastBisectionSteps = 0;
    explicit OCOptimizer(Problem &pr) : p(pr) {}
    void step(double m, double pexp, double ctol, bool inplace) {
        DevVec dJ, dc, x0, stepped, xs, scratch;
        if (inplace) {
            p.evaluateObjectiveGradientDevice(dJ);
        } else {
            auto g = p.evaluateObjectiveGradientAndReturn();
            dJ.upload(g.data(), g.size());
        }
        std::vector<std::unique_ptr<DevVec>> rows;
        p.evaluateConstraintsJacobianDevice(rows);
        if (rows.size() != 1) throw std::runtime_error("Applicable only for a topology optimization with a single (volume) constraint");
        dc.copy_from(*rows[0]);
        x0.copy_from(p.chain->design());
        const size_t n = x0.n;
        stepped.resize(n);
        scratch.resize(n);
        auto ceval = [&](double lambda) {
            check(vf_oc_update((int64_t)n, x0.p, dJ.p, dc.p, lambda, m, pexp, stepped.p, VF_DEVICE, nullptr));
            xs.copy_from(stepped);
            return p.evaluateOCConstraintAtVars(xs, scratch);
        };
        const double dilation = 32;
        double lambda_mid = 0.5 * (lambda_min + lambda_max);
        lambda_max = dilation * lambda_max + (1 - dilation) * lambda_mid;
        lambda_min = std::max(dilation * lambda_min + (1 - dilation) * lambda_mid, 0.01);
        size_t nit = 0;
        for (; nit < loopCountGuard; ++nit) {
            if (ceval(lambda_min) < 0) break;
            lambda_max = lambda_min;
            lambda_min /= 2;
        }
        if (nit == loopCountGuard)
            throw std::runtime_error("Bracketing constraint(lambda_min) < 0 failed (" + std::to_string(loopCountGuard) + " times).");
        if (nit == 0)
            for (; nit < loopCountGuard; ++nit) {
                if (ceval(lambda_max) > 0) break;
                lambda_min = lambda_max;
                lambda_max *= 2;
            }
        if (nit == loopCountGuard)
            throw std::runtime_error("Bracketing constraint(lambda_max) > 0 failed (" + std::to_string(loopCountGuard) + " times).");
        double violation;
        int64_t bis = 0;
        do {
            lambda_mid = 0.5 * (lambda_min + lambda_max);
            violation = ceval(lambda_mid);
            ++bis;
            if (std::abs(violation) <= ctol) break;
            if (violation < 0) lambda_min = lambda_mid;
            if (violation > 0) lambda_max = lambda_mid;
        } while (true);
        lastBisectionSteps = bis;
        p.setVarsFrom(stepped);
    }
};

// ====================================================================================================
// LayerByLayerEvaluator (LayerByLayer.hh:24-311)
// ====================================================================================================
struct LBL {
    TPSPtr tps;
    std::string method = "N=3";
    double total = 0;
    std::vector<double> grad;
    std::vector<double> perLayer;
    std::vector<int64_t> perLayerIters;
    int precision = VF_PREC_FP64;
    // batched evaluation (extension): the layer solves as stacked batches of at most `pipelines` structures
    // (0 = as many as fit) with graph-replayed PCG iterations; batched = false runs the layer solves one
    // after the other on the caller's solver
    bool batched = true;
    int64_t pipelines = 0;
    bool useGraphs = true;
    double seconds = 0;
    int initMethod = VF_LBL_INIT_SUBSPACE;
    int64_t history = 3;
    std::vector<double> uFullDesign;  // full-design solution of the last run (LayerByLayer.hh:282-283)
    explicit LBL(TPSPtr t) : tps(std::move(t)) {}
    // LayerByLayer.hh:215-221
    void selectInitMethod(const std::string &m) {
        if (m == "zero") { initMethod = VF_LBL_INIT_ZERO; history = 0; }
        else if (m == "constant") { initMethod = VF_LBL_INIT_FD; history = 1; }
        else if (m == "fd") { initMethod = VF_LBL_INIT_FD; history = 2; }
        else if (m.rfind("N=", 0) == 0) {
            size_t used = 0;
            long k = 0;
            try { k = std::stol(m.substr(2), &used); } catch (...) { used = 0; }
            if (used == 0 || used != m.size() - 2 || k < 1) throw std::runtime_error("Unrecognized method " + m);
            initMethod = VF_LBL_INIT_SUBSPACE;
            history = k;
        } else {
            throw std::runtime_error("Unrecognized method " + m);
        }
        method = m;
    }
    void run(MG &solver, bool zeroInit, size_t layerIncrement, size_t maxIter, double tol,
             std::function<void(size_t, py::array_t<double>, py::array_t<double>)> it_cb, size_t mgIterations,
             size_t mgSmoothingIterations, bool fmg, bool verbose, py::object lblCallback) {
        const bool callbacks = (bool)it_cb || !lblCallback.is_none();
        const size_t numLayers = (size_t)tps->ne[1];
        const size_t nl = (numLayers + layerIncrement - 1) / layerIncrement;
        grad.assign(tps->nel, 0.0);
        perLayer.assign(nl, 0.0);
        perLayerIters.assign(nl, 0);
        double obj = 0;
        const auto t0 = std::chrono::steady_clock::now();
        // the zero guess makes the layer solves independent, so they can run batched; the other guesses
        // chain each layer to the previous solutions and run one after the other
        const size_t nDof = (size_t)tps->N * tps->nn;
        const bool fullValid = uFullDesign.size() == nDof;
        // callbacks see every iterate and layer in order: they run with the sequential evaluator
        if (batched && !callbacks && initMethod == VF_LBL_INIT_ZERO && (zeroInit || !fullValid)) {
            check(vf_lbl_run_zero_batched(solver.mg, (int64_t)layerIncrement, (int64_t)maxIter, tol,
                                          (int64_t)mgIterations, (int64_t)mgSmoothingIterations, fmg ? 1 : 0, precision,
                                          pipelines, useGraphs ? 1 : 0, &obj, grad.data(), VF_HOST, perLayer.data(),
                                          perLayerIters.data()));
        } else {
            std::vector<double> uf(nDof);
            if (fullValid) uf = uFullDesign;
            struct Ctx {
                std::function<void(size_t, py::array_t<double>, py::array_t<double>)> *it;
                py::object *lbl;
                size_t nn, ne;
                int N;
            } ctx{&it_cb, &lblCallback, (size_t)tps->nn, (size_t)tps->nel, tps->N};
            auto itr = [](void *user, int64_t it, const double *xx, const double *rr) -> int {
                auto *c = (Ctx *)user;
                (*c->it)((size_t)it, soa_to_vfield(xx, c->nn, c->N), soa_to_vfield(rr, c->nn, c->N));
                return 0;
            };
            auto lbr = [](void *user, int64_t l, double compliance, const double *g, const double *u) -> int {
                auto *c = (Ctx *)user;
                (*c->lbl)((size_t)l, compliance, from_vec(std::vector<double>(g, g + c->ne)), soa_to_vfield(u, c->nn, c->N));
                return 0;
            };
            check(vf_lbl_run(solver.mg, (int64_t)layerIncrement, (int64_t)maxIter, tol, (int64_t)mgIterations,
                             (int64_t)mgSmoothingIterations, fmg ? 1 : 0, precision, initMethod, history,
                             zeroInit ? 1 : 0, uf.data(), fullValid ? 1 : 0, VF_HOST, &obj, grad.data(), VF_HOST,
                             perLayer.data(), perLayerIters.data(), it_cb ? +itr : nullptr, &ctx,
                             lblCallback.is_none() ? nullptr : +lbr, &ctx));
            uFullDesign.swap(uf);
        }
        seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total = obj;
        if (verbose) {
            size_t k = 0;
            for (size_t l = numLayers; l > 0; l -= std::min(layerIncrement, l), ++k)
                std::cout << "Layer " << l << ": " << perLayer[k] << std::endl;
        }
    }
};

// ---- boundary conditions from .bc files (MeshFEM readBoundaryConditions as used at
//      TensorProductSimulator.hh:600-658): regions {type, value, box% | box}
void applyBCFile(TPS &t, const std::string &path) {
    if (t.hasBC()) throw std::runtime_error("Boundary condition updates unsupported");
    py::module_ json = py::module_::import("json");
    py::object data;
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Couldn't open boundary condition file " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        data = json.attr("loads")(ss.str());
    }
    const std::array<double, 3> size = {t.hi[0] - t.lo[0], t.hi[1] - t.lo[1], t.hi[2] - t.lo[2]};
    const int N = t.N;  // a planar grid reads the first two entries of every corner and value
    for (auto reg : data["regions"]) {
        auto r = reg.cast<py::dict>();
        const std::string type = r["type"].cast<std::string>();
        std::array<double, 3> bmin, bmax;
        if (r.contains("box%")) {
            auto b = r["box%"].cast<py::dict>();
            auto mn = b["minCorner"].cast<std::vector<double>>(), mx = b["maxCorner"].cast<std::vector<double>>();
            bmin = t.lo; bmax = t.hi;
            for (int d = 0; d < N; ++d) { bmin[d] = t.lo[d] + mn.at(d) * size[d]; bmax[d] = t.lo[d] + mx.at(d) * size[d]; }
        } else if (r.contains("box")) {
            auto b = r["box"].cast<py::dict>();
            auto mn = b["minCorner"].cast<std::vector<double>>(), mx = b["maxCorner"].cast<std::vector<double>>();
            bmin = t.lo; bmax = t.hi;
            for (int d = 0; d < N; ++d) { bmin[d] = mn.at(d); bmax[d] = mx.at(d); }
        } else {
            throw std::runtime_error("Unsupported region specification (need box% or box)");
        }
        auto value = r["value"].cast<std::vector<double>>();
        if ((int)value.size() < N) throw std::runtime_error("BC value must have " + std::to_string(N) + " components");
        value.resize(3, 0.0);
        std::vector<size_t> nodes;
        for (size_t n = 0; n < t.nn; ++n) {
            auto p = t.nodePos(n);
            bool in = true;
            for (int d = 0; d < N; ++d) in &= p[d] >= bmin[d] && p[d] <= bmax[d];
            if (in) nodes.push_back(n);
        }
        if (type == "force") {
            if (nodes.empty()) throw std::runtime_error("Force constraint region unmatched");
            for (size_t n : nodes)
                for (int c = 0; c < N; ++c) t.forces[c * t.nn + n] = value[c] / (double)nodes.size();
        } else if (type.rfind("dirichlet", 0) == 0) {
            if (nodes.empty()) throw std::runtime_error("Dirichlet region unmatched");
            const std::string comps = type.size() > 9 ? type.substr(9) : "xyz";
            uint8_t cm = 0;
            for (char ch : comps) {
                if (ch == 'x') cm |= 1;
                else if (ch == 'y') cm |= 2;
                else if (ch == 'z') cm |= 4;
                else throw std::runtime_error("Illegal constraint type, only \"dirichlet\" and \"force\" accepted");
            }
            for (size_t n : nodes) t.setDirichlet(n, {value[0], value[1], value[2]}, (uint8_t)(cm & t.fullMask()));
        } else {
            throw std::runtime_error("Illegal constraint type, only \"dirichlet\" and \"force\" accepted");
        }
    }
    t.push_bc();
}

uint8_t parseMask(const std::string &s) {
    uint8_t cm = 0;
    for (char ch : s) {
        if (ch == 'x') cm |= 1;
        else if (ch == 'y') cm |= 2;
        else if (ch == 'z') cm |= 4;
        else throw std::runtime_error("Invalid component mask " + s);
    }
    return cm;
}

TPSPtr makeTPS(const std::array<double, 3> &lo, const std::array<double, 3> &hi, const std::array<int64_t, 3> &ne,
               int N = 3) {
    return std::make_shared<TPS>(lo, hi, ne, N);
}

void copyMaterial(const TPS &src, TPS &dst) {
    dst.young = src.young;
    dst.poisson = src.poisson;
    check(vf_sim_set_material(dst.sim, dst.young, dst.poisson));
    // E_0, E_min and the SIMP exponent carry over; the RAMP factor and the law keep the new simulator's
    // defaults (TensorProductSimulator.hh:1876-1884, 1940-1943)
    dst.E0 = src.E0;
    dst.Emin = src.Emin;
    dst.gamma = src.gamma;
}

// getIntermediateFabricationShape (TensorProductSimulator.hh:1864-1923)
// validatedIntermediateFabricationShapeSize (TensorProductSimulator.hh:1990-1999): an intermediate shape has
// this grid's element counts except along the build direction (axis 1), where it has at most as many
void checkIntermediateSize(const TPS &full, const TPS &part) {
    if (part.ne[0] != full.ne[0] || part.ne[2] != full.ne[2] || part.ne[1] > full.ne[1])
        throw std::runtime_error("Intermediate shape is of unexpected size");
}

// isotropic ElasticityTensor value of the ETensor property (MeshFEM's tensor, restricted to the isotropic
// materials this build supports)
struct IsoETensor {
    double E = 1, nu = 0;
    void setIsotropic(double e, double n) { E = e; nu = n; }
};

TPSPtr intermediateShape(const TPS &t, double hfrac, bool validate, InterpolationLaw law) {
    if (hfrac < 0 || hfrac > 1) throw std::runtime_error("hfrac is out of bounds");
    const double fh = hfrac * (double)t.ne[1];
    const int64_t helems = (int64_t)std::llround(fh);
    if (std::abs(fh - (double)helems) > 1e-10) throw std::runtime_error("hfrac chops off a noninteger number of element layers");
    std::array<int64_t, 3> ne = t.ne;
    ne[1] = helems;
    std::array<double, 3> hi = t.hi;
    hi[1] = t.lo[1] + hfrac * (t.hi[1] - t.lo[1]);
    auto r = makeTPS(t.lo, hi, ne, t.N);
    copyMaterial(t, *r);
    r->law = law;
    r->push_law();
    auto rho = t.getDensities();
    std::vector<double> sub(r->nel);
    for (int64_t ex = 0; ex < ne[0]; ++ex)
        for (int64_t ey = 0; ey < ne[1]; ++ey)
            for (int64_t ez = 0; ez < ne[2]; ++ez)
                sub[(ex * ne[1] + ey) * ne[2] + ez] = rho[(ex * t.ne[1] + ey) * t.ne[2] + ez];
    r->setDensities(sub);
    const bool noGravity = t.grav[0] == 0 && t.grav[1] == 0 && t.grav[2] == 0;
    if (noGravity) {
        if (validate) throw std::runtime_error("Original simulator has unexpected boundary conditions for layer-by-layer simulation");
        r->grav = {0, -1, 0};
    } else {
        r->grav = t.grav;
    }
    check(vf_sim_set_gravity(r->sim, r->grav.data()));
    for (size_t n = 0; n < r->nn; ++n)
        if (r->nodePos(n)[1] == r->lo[1]) r->setDirichlet(n, {0, 0, 0}, r->fullMask());
    r->push_bc();
    return r;
}

}  // namespace

// ====================================================================================================
PYBIND11_MODULE(pyVoxelFEM, m) {
    m.doc() = "Voxel-based finite element solve-and-sensitivity path (MI355X HIP backend)";
    py::module_ detail = m.def_submodule("detail");

    py::enum_<InterpolationLaw>(m, "InterpolationLaw")
        .value("SIMP", InterpolationLaw::SIMP)
        .value("RAMP", InterpolationLaw::RAMP)
        .export_values();
    py::enum_<NumberType>(m, "NumberType").value("DOUBLE", NumberType::DOUBLE).value("FLOAT", NumberType::FLOAT);

    // ------------------------------------------------------------------ TensorProductSimulator
    py::class_<IsoETensor>(m, "ElasticityTensor")
        .def(py::init<>())
        .def(py::init([](double E, double nu) { IsoETensor e; e.setIsotropic(E, nu); return e; }), py::arg("E"), py::arg("nu"))
        .def("setIsotropic", &IsoETensor::setIsotropic, py::arg("E"), py::arg("nu"))
        .def_readonly("E", &IsoETensor::E)
        .def_readonly("nu", &IsoETensor::nu)
        .def("__len__", [](const IsoETensor &) { return 2; })
        .def("__getitem__", [](const IsoETensor &e, int i) {
                if (i == 0 || i == -2) return e.E;
                if (i == 1 || i == -1) return e.nu;
                throw py::index_error();
            })
        .def("__repr__", [](const IsoETensor &e) {
                return "ElasticityTensor(isotropic, E=" + std::to_string(e.E) + ", nu=" + std::to_string(e.nu) + ")";
            });

    py::class_<TPS, TPSPtr>(detail, "TensorProductSimulator1_1_1")
        .def("numNodes", [](const TPS &t) { return t.nn; })
        .def("numElements", [](const TPS &t) { return t.nel; })
        .def("getDensities", [](const TPS &t) { return from_vec(t.getDensities()); })
        .def("setDensities", [](TPS &t, const Arr &rho) { t.setDensities(to_vec(rho, t.nel, "rho")); }, py::arg("rho"))
        .def("readMaterial", [](TPS &t, const std::string &path) {
                py::module_ json = py::module_::import("json");
                std::ifstream in(path);
                if (!in) throw std::runtime_error("Couldn't open material file " + path);
                std::stringstream ss; ss << in.rdbuf();
                py::dict d = json.attr("loads")(ss.str());
                const std::string type = d.contains("type") ? d["type"].cast<std::string>() : "isotropic_material";
                if (type != "isotropic_material" && type != "isotropic") throw std::runtime_error("Only isotropic materials are supported");
                t.young = d["young"].cast<double>();
                t.poisson = d["poisson"].cast<double>();
                check(vf_sim_set_material(t.sim, t.young, t.poisson));
            }, py::arg("materialPath"))
        .def("setDensity", [](TPS &t, size_t ei, double v) {
                if (v > 1.0 || v < 0) throw std::runtime_error("Density value (" + std::to_string(v) + ") has to be in between 0 and 1");
                auto r = t.getDensities(); r.at(ei) = v; t.setDensities(r);
            }, py::arg("ei"), py::arg("value"))
        .def("setUniformDensities", [](TPS &t, double v) {
                if (v > 1.0 || v < 0) throw std::runtime_error("Density value (" + std::to_string(v) + ") has to be in between 0 and 1");
                t.setDensities(std::vector<double>(t.nel, v));
            }, py::arg("density"))
        .def("setDensitiesFromCoarseGrid", [](TPS &t, size_t f, const Arr &rho) {
                const size_t nc = (size_t)rho.size();
                if (t.nel != (t.N == 2 ? f * f : f * f * f) * nc) throw std::runtime_error("Incorrect rho size");
                auto r = to_vec(rho, nc, "rho");
                const int64_t c1 = t.ne[1] / f, c2 = t.N == 2 ? 1 : t.ne[2] / f;
                std::vector<double> out(t.nel);
                for (int64_t e = 0; e < (int64_t)t.nel; ++e) {
                    const int64_t ez = e % t.ne[2], ey = (e / t.ne[2]) % t.ne[1], ex = e / (t.ne[1] * t.ne[2]);
                    out[e] = r[((ex / f) * c1 + ey / f) * c2 + (t.N == 2 ? 0 : ez / f)];
                }
                t.setDensities(out);
            }, py::arg("upscalingFactor"), py::arg("rho"))
        .def("elementDensity", [](const TPS &t, size_t e) { return t.getDensities().at(e); }, py::arg("ei"))
        .def("elementYoungModulusScaleFactor", [](const TPS &t, size_t e) { return t.getE().at(e); }, py::arg("ei"))
        .def("getYoungModulusScaleFactor", [](const TPS &t) { return from_vec(t.getE()); })
        .def("setFabricationMaskHeightByLayer", [](TPS &t, int64_t l) { check(vf_sim_set_mask_layer(t.sim, l)); }, py::arg("l"))
        .def("getFabricationMaskHeight", [](const TPS &t) { return t.maskHeight(); })
        .def("solve", [](TPS &t, const Arr &f) {
                auto x = t.solve(vfield_to_soa(f, t.nn, t.N));
                return soa_to_vfield(x.data(), t.nn, t.N);
            }, py::arg("f"))
        .def("solveWithImposedLoads", [](TPS &t) {
                std::vector<double> f((size_t)t.N * t.nn);
                check(vf_sim_build_load(t.sim, f.data(), VF_HOST));
                auto x = t.solve(f);
                return soa_to_vfield(x.data(), t.nn, t.N);
            })
        .def("complianceGradient", [](const TPS &t, const Arr &u) {
                auto us = vfield_to_soa(u, t.nn, t.N);
                std::vector<double> g(t.nel);
                check(vf_sim_compliance_gradient(t.sim, us.data(), g.data(), VF_HOST));
                return from_vec(g);
            }, py::arg("u"))
        .def("multigridSolver", [](TPSPtr t, size_t L) { return std::make_shared<MG>(t, (int64_t)L); }, py::arg("numCoarseningLevels"))
        .def("applyK", [](const TPS &t, const Arr &u) {
                auto us = vfield_to_soa(u, t.nn, t.N);
                std::vector<double> y((size_t)t.N * t.nn);
                check(vf_sim_apply_K(t.sim, us.data(), y.data(), VF_HOST));
                return soa_to_vfield(y.data(), t.nn, t.N);
            }, py::arg("u"))
        .def("getDirichletVarsAndValues", [](const TPS &t) {
                std::vector<size_t> vars; std::vector<double> vals;
                for (size_t n = 0; n < t.nn; ++n)
                    for (int c = 0; c < t.N; ++c)
                        if ((t.dmask[n] >> c) & 1) { vars.push_back(t.N * n + c); vals.push_back(t.dvals[c * t.nn + n]); }
                return std::make_tuple(vars, vals);
            })
        .def("buildLoadVector", [](const TPS &t) {
                std::vector<double> f((size_t)t.N * t.nn);
                check(vf_sim_build_load(t.sim, f.data(), VF_HOST));
                return soa_to_vfield(f.data(), t.nn, t.N);
            })
        .def("nodePosition", [](const TPS &t, size_t n) { return t.dims(t.nodePos(n)); }, py::arg("ni"))
        .def("applyDisplacementsAndLoadsFromFile", [](TPS &t, const std::string &p) { applyBCFile(t, p); }, py::arg("bcPath"))
        .def("elementIndexForGridCell", [](const TPS &t, const std::vector<int64_t> &c) {
                if ((int)c.size() != t.N) throw std::runtime_error("expected " + std::to_string(t.N) + " cell indices");
                for (int d = 0; d < t.N; ++d) if (c[d] < 0 || c[d] >= t.ne[d]) throw std::runtime_error("Indices out of bounds of NDVector");
                return (size_t)((c[0] * t.ne[1] + c[1]) * t.ne[2] + (t.N == 3 ? c[2] : 0));
            }, py::arg("cellIdxs"))
        .def("getForceMask", [](const TPS &t) {
                py::array_t<bool> out({(py::ssize_t)t.nn, (py::ssize_t)t.N});
                auto w = out.mutable_unchecked<2>();
                for (size_t n = 0; n < t.nn; ++n) for (int c = 0; c < t.N; ++c) w(n, c) = t.forces[c * t.nn + n] != 0;
                return out;
            })
        .def("getDirichletMask", [](const TPS &t) {
                py::array_t<bool> out({(py::ssize_t)t.nn, (py::ssize_t)t.N});
                auto w = out.mutable_unchecked<2>();
                for (size_t n = 0; n < t.nn; ++n) for (int c = 0; c < t.N; ++c) w(n, c) = (t.dmask[n] >> c) & 1;
                return out;
            })
        .def("getBCIndicatorField", [](const TPS &t) {
                std::vector<double> r(t.nn, 0.0);
                for (size_t n = 0; n < t.nn; ++n) {
                    r[n] += t.dmask[n] & 7;
                    int fb = 0;
                    for (int c = 0; c < t.N; ++c) if (t.forces[c * t.nn + n] != 0) fb |= 1 << c;
                    r[n] += (t.N == 2 ? 4 : 8) * fb;
                }
                return from_vec(r);
            })
        .def("elementStiffnessMatrix", [](const TPS &t, size_t e) {
                auto K = t.K0(); const double E = t.getE().at(e);
                const int ks = t.keSize();
                py::array_t<double> out({ks, ks});
                auto w = out.mutable_unchecked<2>();
                for (int a = 0; a < ks; ++a) for (int b = 0; b < ks; ++b) w(a, b) = E * K[a * ks + b];
                return out;
            }, py::arg("ei"))
        .def("fullDensityElementStiffnessMatrix", [](const TPS &t) {
                auto K = t.K0();
                const int ks = t.keSize();
                py::array_t<double> out({ks, ks});
                std::memcpy(out.mutable_data(), K.data(), (size_t)ks * ks * sizeof(double));
                return out;
            })
        .def("elementNodes", [](const TPS &t, size_t e) {
                const int nv = t.N == 2 ? 4 : 8;
                std::vector<size_t> r(nv); for (int l = 0; l < nv; ++l) r[l] = t.elemNode(e, l); return r;
            }, py::arg("ei"))
        .def("elemNodeGlobalIndex", [](const TPS &t, size_t e, int l) { return t.elemNode(e, l); }, py::arg("ei"), py::arg("n"))
        .def("addDirichletCondition", [](TPS &t, const std::vector<double> &u, const std::vector<double> &mn,
                                         const std::vector<double> &mx, const std::string &mask) {
                if ((int)u.size() != t.N || (int)mn.size() != t.N || (int)mx.size() != t.N)
                    throw std::runtime_error("expected " + std::to_string(t.N) + "-vectors");
                const uint8_t cm = parseMask(mask) & t.fullMask();
                std::array<double, 3> uv{0, 0, 0};
                for (int d = 0; d < t.N; ++d) uv[d] = u[d];
                for (size_t n = 0; n < t.nn; ++n) {
                    auto p = t.nodePos(n);
                    bool in = true;
                    for (int d = 0; d < t.N; ++d) in &= p[d] >= mn[d] && p[d] <= mx[d];
                    if (in) t.setDirichlet(n, uv, cm);
                }
                t.push_bc();
            }, py::arg("u"), py::arg("minCorner"), py::arg("maxCorner"), py::arg("componentMask") = "xyz")
        .def("applySymmetryConditions", [](TPS &t, const std::vector<bool> &axes, std::vector<bool> face) {
                if ((int)axes.size() != t.N) throw std::runtime_error("expected " + std::to_string(t.N) + " symmetry flags");
                face.resize(t.N, false);
                for (size_t n = 0; n < t.nn; ++n) {
                    auto p = t.nodePos(n);
                    for (int d = 0; d < t.N; ++d) {
                        if (!axes[d]) continue;
                        if (std::abs(p[d] - (face[d] ? t.hi[d] : t.lo[d])) < 1e-10) {
                            t.dmask[n] |= (uint8_t)(1 << d);
                            t.dvals[d * t.nn + n] = 0.0;
                        }
                    }
                }
                t.push_bc();
            }, py::arg("symmetry_axes"), py::arg("minMaxFace") = std::vector<bool>{})
        .def_property_readonly("domain", [](const TPS &t) { return std::make_pair(t.dims(t.lo), t.dims(t.hi)); })
        .def_property_readonly("bbox", [](const TPS &t) { return std::make_pair(t.dims(t.lo), t.dims(t.hi)); })
        .def_property_readonly("gridShape", [](const TPS &t) { return t.dims(t.ne); })
        .def_property_readonly("NbElementsPerDimension", [](const TPS &t) { return t.dims(t.ne); })
        .def_property_readonly("NbNodesPerDimension", [](const TPS &t) { return t.dims(t.nnodes3()); })
        .def_property("interpolationLaw", [](const TPS &t) { return t.law; },
                      [](TPS &t, InterpolationLaw l) { t.law = l; t.push_law(); })
        .def_property("E_0", [](const TPS &t) { return t.E0; }, [](TPS &t, double v) { t.E0 = v; t.push_law(); })
        .def_property("E_min", [](const TPS &t) { return t.Emin; }, [](TPS &t, double v) { t.Emin = v; t.push_law(); })
        .def_property("gamma", [](const TPS &t) { return t.gamma; }, [](TPS &t, double v) { t.gamma = v; t.push_law(); })
        .def_property("q", [](const TPS &t) { return t.q; }, [](TPS &t, double v) { t.q = v; t.push_law(); })
        .def_property("gravity", [](const TPS &t) { return t.dims(t.grav); },
                      [](TPS &t, const std::vector<double> &g) {
                          if ((int)g.size() != t.N) throw std::runtime_error("gravity must have " + std::to_string(t.N) + " components");
                          t.grav = {g[0], g[1], t.N == 3 ? g[2] : 0.0};
                          check(vf_sim_set_gravity(t.sim, t.grav.data()));
                      })
        .def_property("ETensor", [](const TPS &t) { IsoETensor e; e.E = t.young; e.nu = t.poisson; return e; },
                      [](TPS &t, py::object o) {
                          if (py::isinstance<IsoETensor>(o)) {
                              auto e = o.cast<IsoETensor>();
                              t.young = e.E; t.poisson = e.nu;
                          } else {
                              auto ep = o.cast<std::pair<double, double>>();
                              t.young = ep.first; t.poisson = ep.second;
                          }
                          check(vf_sim_set_material(t.sim, t.young, t.poisson));
                      }, "Isotropic elasticity tensor (an ElasticityTensor, or a (Young's modulus, Poisson ratio) pair)")
        .def_property("dx", [](const TPS &t) { return t.dims(t.stretch()); },
                      [](TPS &t, const std::vector<double> &h) {
                          if ((int)h.size() != t.N) throw std::runtime_error("expected " + std::to_string(t.N) + " stretchings");
                          const double hh[3] = {h[0], h[1], t.N == 3 ? h[2] : 1.0};
                          check(vf_sim_set_stretchings(t.sim, hh));
                      },
                      "Get the dimensions of a grid cell")
        .def_property_readonly("elementVolume", [](const TPS &t) {
                auto h = t.stretch();
                return t.N == 2 ? h[0] * h[1] : h[0] * h[1] * h[2];
            })
        .def("elementEnergyDensity", [](const TPS &t, const Arr &u) {
                auto us = vfield_to_soa(u, t.nn, t.N);
                std::vector<double> out(t.nel);
                check(vf_sim_energy_density(t.sim, us.data(), out.data(), VF_HOST));
                return from_vec(out);
            }, py::arg("u"))
        .def("getK", [](const TPS &t) {
                // getK (TensorProductSimulator.hh:1506-1513, assembly 833-865): E_e K0 summed into the upper
                // triangle (row <= column, DoF 3 node + component) of the stiffness matrix, as a scipy CSC matrix
                const auto K = t.K0();
                const auto E = t.getE();
                const int ks = t.keSize(), nv = ks / t.N;
                const size_t nt = t.nel * (size_t)(ks * (ks + 1) / 2);
                py::array_t<double> vals((py::ssize_t)nt);
                py::array_t<int64_t> rows((py::ssize_t)nt), cols((py::ssize_t)nt);
                double *v = vals.mutable_data();
                int64_t *ri = rows.mutable_data(), *ci = cols.mutable_data();
                size_t k = 0;
                for (size_t e = 0; e < t.nel; ++e) {
                    size_t dofs[24];
                    for (int l = 0; l < nv; ++l)
                        for (int c = 0; c < t.N; ++c) dofs[t.N * l + c] = t.N * t.elemNode(e, l) + c;
                    for (int a = 0; a < ks; ++a)
                        for (int b = 0; b < ks; ++b) {
                            if (dofs[a] > dofs[b]) continue;
                            v[k] = E[e] * K[a * ks + b];
                            ri[k] = (int64_t)dofs[a];
                            ci[k] = (int64_t)dofs[b];
                            ++k;
                        }
                }
                py::module_ sp = py::module_::import("scipy.sparse");
                const py::ssize_t n = (py::ssize_t)(t.N * t.nn);
                return sp.attr("csc_matrix")(py::make_tuple(vals, py::make_tuple(rows, cols)), py::make_tuple(n, n));
            })
        .def("constantStrainLoad", [](const TPS &t, const Arr &eps) {
                // the strain as a 3x3 symmetric matrix or as the 6 values (xx, yy, zz, yz, xz, xy); on a planar
                // grid a 2x2 matrix or the 3 values (xx, yy, xy)
                double e[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (t.N == 2) {
                    if (eps.ndim() == 2 && eps.shape(0) == 2 && eps.shape(1) == 2) {
                        auto r = eps.unchecked<2>();
                        for (int a = 0; a < 2; ++a) for (int b = 0; b < 2; ++b) e[2 * a + b] = r(a, b);
                    } else if (eps.size() == 3) {
                        const double *q = eps.data();
                        e[0] = q[0]; e[3] = q[1]; e[1] = e[2] = q[2];
                    } else {
                        throw std::runtime_error("constantStrainLoad: eps must be a 2x2 symmetric matrix or 3 values");
                    }
                } else if (eps.ndim() == 2 && eps.shape(0) == 3 && eps.shape(1) == 3) {
                    auto r = eps.unchecked<2>();
                    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) e[3 * a + b] = r(a, b);
                } else if (eps.size() == 6) {
                    const double *q = eps.data();
                    e[0] = q[0]; e[4] = q[1]; e[8] = q[2];
                    e[5] = e[7] = q[3];
                    e[2] = e[6] = q[4];
                    e[1] = e[3] = q[5];
                } else {
                    throw std::runtime_error("constantStrainLoad: eps must be a 3x3 symmetric matrix or 6 values");
                }
                std::vector<double> F((size_t)t.N * t.nn);
                check(vf_sim_constant_strain_load(t.sim, e, F.data(), VF_HOST));
                return soa_to_vfield(F.data(), t.nn, t.N);
            }, py::arg("eps"))
        .def("sampleNodalField", [](TPS &t, const Arr &u, const Arr &p) {
                if (u.ndim() < 1 || (size_t)u.shape(0) != t.nn) throw std::runtime_error("u must have one row per node");
                if (p.ndim() != 2 || p.shape(1) != t.N)
                    throw std::runtime_error("p must hold one " + std::to_string(t.N) + "D point per row");
                const int64_t k = u.ndim() == 1 ? 1 : (int64_t)u.shape(1);
                const int64_t np = (int64_t)p.shape(0);
                std::vector<double> us((size_t)k * t.nn), ps((size_t)t.N * np), out((size_t)k * np);
                auto uu = u.unchecked();
                for (size_t n = 0; n < t.nn; ++n)
                    for (int64_t c = 0; c < k; ++c)
                        us[(size_t)c * t.nn + n] = u.ndim() == 1 ? u.data()[n] : u.at(n, c);
                (void)uu;
                std::memcpy(ps.data(), p.data(), ps.size() * sizeof(double));
                check(vf_sim_sample_nodal_field(t.sim, us.data(), k, ps.data(), np, out.data(), VF_HOST));
                py::array_t<double> res({(py::ssize_t)np, (py::ssize_t)k});
                auto w = res.mutable_unchecked<2>();
                for (int64_t i = 0; i < np; ++i)
                    for (int64_t c = 0; c < k; ++c) w(i, c) = out[(size_t)c * np + i];
                return res;
            }, py::arg("u"), py::arg("p"), "Sample the nodal FEM field `u` at the query points specified by the rows of `p`")
        .def("transferDensitiesToIntermediateFabricationShape", [](const TPS &t, TPS &part) {
                checkIntermediateSize(t, part);
                const auto rho = t.getDensities();
                std::vector<double> r(part.nel);
                for (size_t e = 0; e < part.nel; ++e) {
                    const size_t ez = e % part.ne[2], ey = (e / part.ne[2]) % part.ne[1], ex = e / (part.ne[1] * part.ne[2]);
                    r[e] = rho[(ex * t.ne[1] + ey) * t.ne[2] + ez];
                }
                part.setDensities(r);
            }, py::arg("intermediateTPS"))
        .def("transferVFieldToIntermediateFabricationShape", [](const TPS &t, const TPS &part, const Arr &u) {
                checkIntermediateSize(t, part);
                if (u.ndim() != 2 || (size_t)u.shape(0) != t.nn || u.shape(1) != t.N)
                    throw std::runtime_error("expected a nodal field of shape (" + std::to_string(t.nn) + ", " + std::to_string(t.N) + ")");
                const auto N = t.nnodes3(), P = part.nnodes3();
                py::array_t<double> res({(py::ssize_t)part.nn, (py::ssize_t)t.N});
                auto w = res.mutable_unchecked<2>();
                auto r = u.unchecked<2>();
                for (int64_t ix = 0; ix < P[0]; ++ix)
                    for (int64_t iy = 0; iy < P[1]; ++iy)
                        for (int64_t iz = 0; iz < P[2]; ++iz) {
                            const size_t src = (ix * N[1] + iy) * N[2] + iz, dst = (ix * P[1] + iy) * P[2] + iz;
                            for (int c = 0; c < t.N; ++c) w(dst, c) = r(src, c);
                        }
                return res;
            }, py::arg("intermediateTPS"), py::arg("u"))
        .def("accumElementScalarFieldFromIntermediateFabricationShape",
             [](const TPS &t, const TPS &part, const Arr &rho_in, py::array_t<double> rho_accum) {
                 // rho_accum is updated in place (the reference binds it as an Eigen::Ref)
                 checkIntermediateSize(t, part);
                 if ((size_t)rho_in.size() != part.nel) throw std::runtime_error("rho_in size mismatch");
                 if ((size_t)rho_accum.size() != t.nel || rho_accum.ndim() != 1 || !rho_accum.writeable() ||
                     rho_accum.strides(0) != (py::ssize_t)sizeof(double))
                     throw std::runtime_error("rho_accum must be a writable contiguous float64 vector of numElements values");
                 double *acc = rho_accum.mutable_data();
                 const double *in = rho_in.data();
                 for (size_t e = 0; e < part.nel; ++e) {
                     const size_t ez = e % part.ne[2], ey = (e / part.ne[2]) % part.ne[1], ex = e / (part.ne[1] * part.ne[2]);
                     acc[(ex * t.ne[1] + ey) * t.ne[2] + ez] += in[e];
                 }
             }, py::arg("intermediateTPS"), py::arg("rho_in"), py::arg("rho_accum"))
        .def("getMesh", [](const TPS &t) {
                // getMesh (TensorProductSimulator.hh:747-776): node positions and hexahedra in Gmsh node order
                py::array_t<double> V({(py::ssize_t)t.nn, (py::ssize_t)t.N});
                auto v = V.mutable_unchecked<2>();
                for (size_t n = 0; n < t.nn; ++n) {
                    auto pos = t.nodePos(n);
                    for (int c = 0; c < t.N; ++c) v(n, c) = pos[c];
                }
                if (t.N == 2) {
                    // quads, counter-clockwise from the lower-left corner (local nodes 0, 2, 3, 1)
                    py::array_t<int64_t> Q({(py::ssize_t)t.nel, (py::ssize_t)4});
                    auto q = Q.mutable_unchecked<2>();
                    const int order[4] = {0, 2, 3, 1};
                    for (size_t e = 0; e < t.nel; ++e)
                        for (int k = 0; k < 4; ++k) q(e, k) = (int64_t)t.elemNode(e, order[k]);
                    return py::make_tuple(V, Q);
                }
                py::array_t<int64_t> F({(py::ssize_t)t.nel, (py::ssize_t)8});
                auto f = F.mutable_unchecked<2>();
                for (size_t e = 0; e < t.nel; ++e)
                    for (int pair = 0; pair < 4; ++pair) {
                        const int64_t v1 = (int64_t)t.elemNode(e, 2 * pair), v2 = (int64_t)t.elemNode(e, 2 * pair + 1);
                        f(e, 2 * pair) = pair % 2 == 0 ? v1 : v2;
                        f(e, 2 * pair + 1) = pair % 2 == 0 ? v2 : v1;
                    }
                return py::make_tuple(V, F);
            })
        .def("mesh", [](TPSPtr t) { return t; }, "The simulator itself (as the reference, for mesh helpers)")
        .def("clearCachedElementStiffness", [](TPS &) {
                // no custom per-element stiffness is cached by this build: every element uses E_e K0
            })
        .def("debugMulticolorElementVisit", [](const TPS &t) {
                // visitElementsMulticolored (TensorProductSimulator.hh:1444-1457) order, serial: colour offsets
                // (x, y, z parities, x slowest), then the elements of a colour in C order
                std::vector<double> r(t.nel, 0.0);
                double i = 0;
                for (int colour = 0; colour < 8; ++colour) {
                    if (t.N == 2 && (colour & 1)) continue;  // planar: z parity 0 only (ne[2] = 1)
                    const int cx = (colour >> 2) & 1, cy = (colour >> 1) & 1, cz = colour & 1;
                    for (int64_t x = cx; x < t.ne[0]; x += 2)
                        for (int64_t y = cy; y < t.ne[1]; y += 2)
                            for (int64_t z = cz; z < t.ne[2]; z += 2) r[(x * t.ne[1] + y) * t.ne[2] + z] = i++;
                }
                return from_vec(r);
            })
        .def("getIntermediateFabricationShape", [](const TPS &t, double yfrac, bool validate, InterpolationLaw law) {
                return intermediateShape(t, yfrac, validate, law);
            }, py::arg("yfrac"), py::arg("validateBoundaryConditions") = true, py::arg("law") = InterpolationLaw::SIMP)
        .def("downsample", [](const TPS &t, size_t levels) {
                const int64_t f = (int64_t)1 << levels;
                std::array<int64_t, 3> ne = t.ne;
                for (int d = 0; d < t.N; ++d) {
                    if (ne[d] % f) throw std::runtime_error("Grid size must be divisible by 2^downsamplingLevels");
                    ne[d] /= f;
                }
                auto r = makeTPS(t.lo, t.hi, ne, t.N);
                copyMaterial(t, *r);
                r->push_law();
                return r;
            }, py::arg("downsamplingLevels"))
        .def("downsampleDensityFieldTo", [](const TPS &t, const Arr &rho, TPS &c) {
                const int64_t f = t.ne[0] / c.ne[0];
                for (int d = 0; d < t.N; ++d) if (f * c.ne[d] != t.ne[d]) throw std::runtime_error("Invalid downsampled simulator");
                auto r = to_vec(rho, t.nel, "input densities");
                std::vector<double> out(c.nel, 0.0);
                const double w = 1.0 / (double)(t.N == 2 ? f * f : f * f * f);
                for (int64_t e = 0; e < (int64_t)t.nel; ++e) {
                    const int64_t ez = e % t.ne[2], ey = (e / t.ne[2]) % t.ne[1], ex = e / (t.ne[1] * t.ne[2]);
                    out[((ex / f) * c.ne[1] + ey / f) * c.ne[2] + (t.N == 2 ? 0 : ez / f)] += r[e] * w;
                }
                c.setDensities(out);
            }, py::arg("densities"), py::arg("coarseTPS"))
        .def("upsampleDensityGradientFrom", [](const TPS &t, const TPS &c, const Arr &g) {
                const int64_t f = t.ne[0] / c.ne[0];
                for (int d = 0; d < t.N; ++d) if (f * c.ne[d] != t.ne[d]) throw std::runtime_error("Invalid downsampled simulator");
                auto gc = to_vec(g, c.nel, "coarse gradient");
                std::vector<double> out(t.nel);
                const double w = 1.0 / (double)(t.N == 2 ? f * f : f * f * f);
                for (int64_t e = 0; e < (int64_t)t.nel; ++e) {
                    const int64_t ez = e % t.ne[2], ey = (e / t.ne[2]) % t.ne[1], ex = e / (t.ne[1] * t.ne[2]);
                    out[e] = gc[((ex / f) * c.ne[1] + ey / f) * c.ne[2] + (t.N == 2 ? 0 : ez / f)] * w;
                }
                return from_vec(out);
            }, py::arg("coarseTPS"), py::arg("g_coarse"))
        // extension: the C ABI handle (vf_sim*) as an integer, for device-resident callers
        .def("_handle", [](const TPS &t) { return (uintptr_t)t.sim; })
        // extension (no reference counterpart): nodal point forces set directly, (numNodes, 3)
        .def("setForces", [](TPS &t, const Arr &f) { t.forces = vfield_to_soa(f, t.nn, t.N); t.push_bc(); }, py::arg("f"))
        // extension: Dirichlet conditions from a (numNodes, 3) boolean mask (zero displacement)
        .def("setDirichletMask", [](TPS &t, py::array_t<bool, py::array::forcecast> mask) {
                if (mask.ndim() != 2 || (size_t)mask.shape(0) != t.nn || mask.shape(1) != t.N)
                    throw std::runtime_error("expected a (numNodes, " + std::to_string(t.N) + ") mask");
                auto r = mask.unchecked<2>();
                for (size_t n = 0; n < t.nn; ++n) {
                    uint8_t m = 0;
                    for (int c = 0; c < t.N; ++c) m |= r(n, c) ? (uint8_t)(1 << c) : 0;
                    t.dmask[n] = m;
                    for (int c = 0; c < t.N; ++c) t.dvals[c * t.nn + n] = 0.0;
                }
                t.push_bc();
            }, py::arg("mask"));

    py::class_<LevelView>(detail, "MultigridLevel")
        .def("numNodes", [](const LevelView &v) { return v.owner->nodes(v.l); })
        .def_property_readonly("NbNodesPerDimension", [](const LevelView &v) {
                int64_t n3[3]; check(vf_mg_level_nodes(v.owner->mg, v.l, n3));
                return std::vector<int64_t>(n3, n3 + v.owner->tps->N);
            })
        .def_property_readonly("NbElementsPerDimension", [](const LevelView &v) {
                int64_t n3[3]; check(vf_mg_level_nodes(v.owner->mg, v.l, n3));
                std::vector<int64_t> r;
                for (int d = 0; d < v.owner->tps->N; ++d) r.push_back(n3[d] - 1);
                return r;
            })
        .def("getDirichletMask", [](const LevelView &v) {
                const size_t nn = v.owner->nodes(v.l);
                std::vector<uint8_t> m(nn);
                check(vf_mg_level_dirichlet(v.owner->mg, v.l, m.data()));
                const int N = v.owner->tps->N;
                py::array_t<bool> out({(py::ssize_t)nn, (py::ssize_t)N});
                auto w = out.mutable_unchecked<2>();
                for (size_t n = 0; n < nn; ++n) for (int c = 0; c < N; ++c) w(n, c) = (m[n] >> c) & 1;
                return out;
            });

    // ------------------------------------------------------------------ MultigridSolver
    py::class_<MG, MGPtr>(detail, "MultigridSolver1_1_1")
        .def("getSimulator", [](MGPtr mg, size_t l) -> py::object {
                if (l == 0) return py::cast(mg->tps);
                if ((int64_t)l > mg->levels) throw py::index_error("level out of range");
                return py::cast(LevelView{mg, (int64_t)l});
            }, py::arg("l"))
        .def("computeResidual", [](MG &mg, size_t l, const Arr &u, const Arr &b) {
                const size_t nn = mg.nodes(l);
                auto us = vfield_to_soa(u, nn, mg.tps->N), bs = vfield_to_soa(b, nn, mg.tps->N);
                std::vector<double> r((size_t)mg.tps->N * nn);
                check(vf_mg_residual(mg.mg, l, us.data(), bs.data(), r.data(), VF_HOST));
                return soa_to_vfield(r.data(), nn, mg.tps->N);
            }, py::arg("l"), py::arg("u"), py::arg("b"))
        .def("applyK", [](MG &mg, size_t l, const Arr &u) {
                const size_t nn = mg.nodes(l);
                auto us = vfield_to_soa(u, nn, mg.tps->N);
                std::vector<double> y((size_t)mg.tps->N * nn);
                check(vf_mg_apply_K(mg.mg, l, us.data(), y.data(), VF_HOST));
                return soa_to_vfield(y.data(), nn, mg.tps->N);
            }, py::arg("l"), py::arg("u"))
        .def("zeroOutDirichletComponents", [](MG &mg, size_t l, const Arr &u) {
                const size_t nn = mg.nodes(l);
                auto us = vfield_to_soa(u, nn, mg.tps->N);
                check(vf_mg_zero_dirichlet(mg.mg, l, us.data(), VF_HOST));
                return soa_to_vfield(us.data(), nn, mg.tps->N);
            }, py::arg("l"), py::arg("u"))
        .def("updateStiffnessMatrices", [](MG &mg) {
            check(vf_mg_update_stiffness(mg.mg));
            check(vf_mg_finish_update(mg.mg));  // failures reported here, as the reference does
        })
        .def("setSymmetricGaussSeidel", [](MG &mg, bool s) { check(vf_mg_set_symmetric_gs(mg.mg, s ? 1 : 0)); }, py::arg("symmetric"))
        .def("solve", [](MG &mg, const Arr &u, const Arr &f, size_t numSteps, size_t nu, bool stiffnessUpdated,
                         bool zeroDirichlet, std::function<void(size_t, py::array_t<double>)> cb, bool fmg) {
                const size_t nn = mg.tps->nn;
                auto x = vfield_to_soa(u, nn, mg.tps->N), fs = vfield_to_soa(f, nn, mg.tps->N);
                struct Ctx { std::function<void(size_t, py::array_t<double>)> *cb; size_t nn; int N; } ctx{&cb, nn, mg.tps->N};
                auto tr = [](void *user, int64_t it, const double *xx) -> int {
                    auto *c = (Ctx *)user;
                    (*c->cb)((size_t)it, soa_to_vfield(xx, c->nn, c->N));
                    return 0;
                };
                check(vf_mg_solve(mg.mg, x.data(), fs.data(), numSteps, nu, stiffnessUpdated ? 1 : 0, zeroDirichlet ? 1 : 0,
                                  fmg ? 1 : 0, cb ? +tr : nullptr, &ctx, VF_HOST));
                return soa_to_vfield(x.data(), nn, mg.tps->N);
            }, py::arg("u"), py::arg("f"), py::arg("numSteps"), py::arg("numSmoothingSteps"),
            py::arg("stiffnessUpdated") = false, py::arg("zeroDirichlet") = false, py::arg("it_callback") = nullptr,
            py::arg("fullMultigrid") = false)
        .def_property("useGraphs", [](const MG &mg) { return mg.useGraphs; },
                      [](MG &mg, bool v) { check(vf_mg_set_use_graphs(mg.mg, v ? 1 : 0)); mg.useGraphs = v; })
        .def("setFabricationMaskHeightByLayer", [](MG &mg, size_t h) { check(vf_mg_set_mask_layer(mg.mg, (int64_t)h)); }, py::arg("h"))
        .def_property("partialStiffnessUpdates", [](const MG &mg) { return mg.partialUpdates; },
                      [](MG &mg, bool on) { check(vf_mg_set_partial_updates(mg.mg, on ? 1 : 0)); mg.partialUpdates = on; },
                      "Recompute only the band of coarse rows touched by removed layers (the reference's partial update)")
        .def_property_readonly("partialUpdateCount", [](const MG &mg) {
                int64_t c = 0;
                check(vf_mg_partial_update_count(mg.mg, &c));
                return c;
            })
        .def("preconditionedConjugateGradient",
             [](MG &mg, const Arr &u, const Arr &b, size_t maxIter, double tol,
                std::function<void(size_t, py::array_t<double>, py::array_t<double>)> cb, size_t mgIterations,
                size_t mgSmoothingIterations, bool fmg, const std::string &precision) {
                const size_t nn = mg.tps->nn;
                auto x = vfield_to_soa(u, nn, mg.tps->N), bs = vfield_to_soa(b, nn, mg.tps->N);
                int prec;
                if (precision == "fp64" || precision == "double") prec = VF_PREC_FP64;
                else if (precision == "fp32" || precision == "float" || precision == "mixed") prec = VF_PREC_FP32;
                else throw std::runtime_error("precision must be 'fp64' or 'fp32'");
                struct Ctx { std::function<void(size_t, py::array_t<double>, py::array_t<double>)> *cb; size_t nn; int N; } ctx{&cb, nn, mg.tps->N};
                auto tr = [](void *user, int64_t it, const double *xx, const double *rr) -> int {
                    auto *c = (Ctx *)user;
                    (*c->cb)((size_t)it, soa_to_vfield(xx, c->nn, c->N), soa_to_vfield(rr, c->nn, c->N));
                    return 0;
                };
                vf_pcg_stats st{};
                check(vf_mg_pcg(mg.mg, x.data(), bs.data(), maxIter, tol, mgIterations, mgSmoothingIterations, fmg ? 1 : 0,
                                prec, 0, cb ? +tr : nullptr, &ctx, VF_HOST, &st));
                return soa_to_vfield(x.data(), nn, mg.tps->N);
            },
             py::arg("u"), py::arg("b"), py::arg("maxIter"), py::arg("tol"), py::arg("it_callback") = nullptr,
             py::arg("mgIterations") = 1, py::arg("mgSmoothingIterations") = 1, py::arg("fullMultigrid") = false,
             py::arg("precision") = "fp64")
        .def("_handle", [](const MG &mg) { return (uintptr_t)mg.mg; })
        // direct access to single hierarchy operations (C ABI), used by parity tests and drivers
        .def("_smooth", [](MG &mg, size_t l, const Arr &u, const Arr &b, bool forward) {
                const size_t nn = mg.nodes(l);
                auto us = vfield_to_soa(u, nn, mg.tps->N), bs = vfield_to_soa(b, nn, mg.tps->N);
                check(vf_mg_smooth(mg.mg, l, us.data(), bs.data(), forward ? 1 : 0, VF_HOST));
                return soa_to_vfield(us.data(), nn, mg.tps->N);
            }, py::arg("l"), py::arg("u"), py::arg("b"), py::arg("forward") = true)
        .def("_restrict", [](MG &mg, size_t l, const Arr &r) {
                const size_t nf = mg.nodes(l), nc = mg.nodes(l + 1);
                auto rs = vfield_to_soa(r, nf, mg.tps->N);
                std::vector<double> out((size_t)mg.tps->N * nc);
                check(vf_mg_restrict(mg.mg, l, rs.data(), out.data(), VF_HOST));
                return soa_to_vfield(out.data(), nc, mg.tps->N);
            }, py::arg("l"), py::arg("r"))
        .def("_interpolate", [](MG &mg, size_t l, const Arr &xc, const Arr &xf, bool accumulate) {
                const size_t nf = mg.nodes(l), nc = mg.nodes(l + 1);
                auto cs = vfield_to_soa(xc, nc, mg.tps->N), fs = vfield_to_soa(xf, nf, mg.tps->N);
                check(vf_mg_interpolate(mg.mg, l, cs.data(), fs.data(), accumulate ? 1 : 0, VF_HOST));
                return soa_to_vfield(fs.data(), nf, mg.tps->N);
            }, py::arg("l"), py::arg("xc"), py::arg("xf"), py::arg("accumulate") = false)
        .def("_coarseSolve", [](MG &mg, const Arr &b) {
                const size_t nc = mg.nodes(mg.levels);
                auto bs = vfield_to_soa(b, nc, mg.tps->N);
                std::vector<double> x((size_t)mg.tps->N * nc);
                check(vf_mg_coarse_solve(mg.mg, bs.data(), x.data(), VF_HOST));
                return soa_to_vfield(x.data(), nc, mg.tps->N);
            }, py::arg("b"))
        .def("_precondition", [](MG &mg, const Arr &r, size_t mgIterations, size_t nu, bool fmg, const std::string &prec) {
                const size_t nn = mg.tps->nn;
                auto rs = vfield_to_soa(r, nn, mg.tps->N);
                std::vector<double> out((size_t)mg.tps->N * nn);
                check(vf_mg_precondition(mg.mg, rs.data(), out.data(), mgIterations, nu, fmg ? 1 : 0,
                                         prec == "fp32" ? VF_PREC_FP32 : VF_PREC_FP64, VF_HOST));
                return soa_to_vfield(out.data(), nn, mg.tps->N);
            }, py::arg("r"), py::arg("mgIterations") = 1, py::arg("nu") = 1, py::arg("fullMultigrid") = false,
            py::arg("precision") = "fp64")
        .def("debug_get_x", [](MG &mg, int l) {
                const size_t nn = mg.nodes(l);
                std::vector<double> v((size_t)mg.tps->N * nn);
                check(vf_mg_debug_get(mg.mg, l, 0, v.data(), VF_HOST));
                return soa_to_vfield(v.data(), nn, mg.tps->N);
            }, py::arg("l"))
        .def("debug_get_b", [](MG &mg, int l) {
                const size_t nn = mg.nodes(l);
                std::vector<double> v((size_t)mg.tps->N * nn);
                check(vf_mg_debug_get(mg.mg, l, 1, v.data(), VF_HOST));
                return soa_to_vfield(v.data(), nn, mg.tps->N);
            }, py::arg("l"))
        .def("debugMulticolorVisit", [](MG &mg) {
                auto &t = *mg.tps;
                const auto N = t.nnodes3();
                std::vector<int> r(t.nn, 0);
                int i = 0;
                for (int lni = 0; lni < 8; ++lni) {
                    const int cx = (lni >> 2) & 1, cy = (lni >> 1) & 1, cz = lni & 1;
                    for (int64_t x = cx; x < N[0]; x += 2)
                        for (int64_t y = cy; y < N[1]; y += 2)
                            for (int64_t z = cz; z < N[2]; z += 2) r[(x * N[1] + y) * N[2] + z] = i++;
                }
                return r;
            });

    // ------------------------------------------------------------------ filters and constraints
    py::class_<Filter, FilterPtr>(detail, "Filter")
        .def("setInputDimensions", [](Filter &f, const std::vector<int64_t> &d) { f.setInputDims(d); }, py::arg("gridDims"))
        .def("setOutputDimensions", [](Filter &f, const std::vector<int64_t> &d) { f.setOutputDims(d); }, py::arg("gridDims"))
        .def_property_readonly("inputDimensions", [](const Filter &f) { return f.inDims; })
        .def_property_readonly("outputDimensions", [](const Filter &f) { return f.outDims; })
        .def("apply", [](Filter &f, const Arr &x) {
                f.checkDims();
                DevVec in, out;
                auto v = to_vec(x, f.inSize(), "input");
                in.upload(v.data(), v.size());
                f.apply(in, out);
                return from_vec(out.host());
            }, py::arg("x"));

    py::class_<FilterChain, FilterChainPtr>(m, "FilterChain")
        .def(py::init<std::vector<FilterPtr>, const std::vector<int64_t> &>(), py::arg("filters"), py::arg("outGridDimensions"))
        .def("numVars", &FilterChain::size)
        .def("numPhysicalVars", &FilterChain::size)  // the reference binds numVars here as well (VoxelFEM.cc:331)
        .def("gridDims", [](const FilterChain &c) { return c.inDims; }, "Input grid dimensions")
        .def("physicalGridDims", [](const FilterChain &c) { return c.outDims; }, "Output grid dimensions")
        .def("setInputDimensions", &FilterChain::setInputDimensions, py::arg("dims"))
        .def("setOutputDimensions", &FilterChain::setOutputDimensions, py::arg("dims"))
        .def("setDesignVars", [](FilterChain &c, const Arr &x) {
                auto v = to_vec(x, c.size(), "design variable");
                DevVec d; d.upload(v.data(), v.size());
                c.setDesignVarsDevice(d);
            }, py::arg("xDesign"))
        .def("backprop", [](FilterChain &c, const Arr &g) {
                if ((size_t)g.size() != c.numPhysicalVars()) throw std::runtime_error("Size mismatch");
                DevVec d; d.upload(g.data(), c.numPhysicalVars());
                c.backprop(d);
                return from_vec(d.host());
            }, py::arg("g"))
        .def("designVars", [](FilterChain &c) { return from_vec(c.design().host()); })
        .def("physicalVars", [](FilterChain &c) { return from_vec(c.physical().host()); })
        .def_property_readonly("filters", [](const FilterChain &c) { return c.filters; });

    py::class_<PythonFilter, Filter, std::shared_ptr<PythonFilter>>(m, "PythonFilter")
        .def(py::init<>())
        .def_readwrite("apply_cb", &PythonFilter::apply_cb)
        .def_readwrite("backprop_cb", &PythonFilter::backprop_cb);

    py::class_<ProjectionFilter, Filter, std::shared_ptr<ProjectionFilter>>(m, "ProjectionFilter")
        .def(py::init<double>(), py::arg("beta"))
        .def(py::init<>())
        .def("invert", &ProjectionFilter::invert, py::arg("filteredValue"))
        .def_property("beta", [](const ProjectionFilter &p) { return p.beta; }, [](ProjectionFilter &p, double b) {
                if (b <= 0) throw std::runtime_error("Beta parameter has to be positive (received beta = " + std::to_string(b) + ")");
                p.beta = b;
            });

    py::class_<SmoothingFilter, Filter, std::shared_ptr<SmoothingFilter>> pySF(m, "SmoothingFilter");
    py::enum_<SmoothingFilter::Type>(pySF, "Type")
        .value("Const", SmoothingFilter::Type::Const)
        .value("Linear", SmoothingFilter::Type::Linear);
    pySF.def(py::init<int, SmoothingFilter::Type>(), py::arg("radius") = 1, py::arg("type") = SmoothingFilter::Type::Const)
        .def_readwrite("radius", &SmoothingFilter::radius)
        .def_readwrite("type", &SmoothingFilter::type)
        // extension: real-valued hat radius (w = rmin - |d| > 0); 0 restores the integer convention
        .def_readwrite("rmin", &SmoothingFilter::rmin);

    py::class_<UpsampleFilter, Filter, std::shared_ptr<UpsampleFilter>>(m, "UpsampleFilter")
        .def(py::init<int64_t>(), py::arg("factor"))
        .def_readonly("factor", &UpsampleFilter::factor);
    py::class_<VertexToCellFilter, Filter, std::shared_ptr<VertexToCellFilter>>(m, "VertexToCellFilter")
        .def(py::init<>());
    py::class_<LangelaarFilter, Filter, std::shared_ptr<LangelaarFilter>>(m, "LangelaarFilter")
        .def(py::init<>());

    py::class_<Constraint, ConstraintPtr>(detail, "Constraint");
    py::class_<TotalVolumeConstraint, Constraint, std::shared_ptr<TotalVolumeConstraint>>(m, "TotalVolumeConstraint")
        .def(py::init<double>(), py::arg("volumeFraction"))
        .def_readwrite("volumeFraction", &TotalVolumeConstraint::volumeFraction);

    // ------------------------------------------------------------------ objectives
    py::class_<Objective, ObjectivePtr>(detail, "Objective1_1_1");
    py::class_<ComplianceObjective, Objective, std::shared_ptr<ComplianceObjective>>(detail, "ComplianceObjective1_1_1")
        .def("gradient", [](ComplianceObjective &o) { DevVec g; o.gradient(g); return from_vec(g.host()); })
        .def("compliance", &ComplianceObjective::compliance)
        .def("u", [](ComplianceObjective &o) { auto h = o.u.host(); return soa_to_vfield(h.data(), o.tps->nn, o.tps->N); })
        .def("f", [](ComplianceObjective &o) { auto h = o.f.host(); return soa_to_vfield(h.data(), o.tps->nn, o.tps->N); });
    py::class_<MultigridComplianceObjective, ComplianceObjective, std::shared_ptr<MultigridComplianceObjective>>(
        detail, "MultigridComplianceObjective1_1_1")
        .def_readonly("mg", &MultigridComplianceObjective::mg)
        .def_readwrite("cgIter", &MultigridComplianceObjective::cgIter)
        .def_readwrite("tol", &MultigridComplianceObjective::tol)
        .def_readwrite("mgIterations", &MultigridComplianceObjective::mgIterations)
        .def_readwrite("mgSmoothingIterations", &MultigridComplianceObjective::mgSmoothingIterations)
        .def_readwrite("fullMultigrid", &MultigridComplianceObjective::fullMultigrid)
        .def_readwrite("zeroInit", &MultigridComplianceObjective::zeroInit)
        .def_readwrite("residual_cb", &MultigridComplianceObjective::residual_cb)
        .def_property("precision",
                      [](const MultigridComplianceObjective &o) { return o.precision == VF_PREC_FP32 ? "fp32" : "fp64"; },
                      [](MultigridComplianceObjective &o, const std::string &p) {
                          if (p == "fp64") o.precision = VF_PREC_FP64;
                          else if (p == "fp32") o.precision = VF_PREC_FP32;
                          else throw std::runtime_error("precision must be 'fp64' or 'fp32'");
                      })
        .def_readonly("lastIterations", &MultigridComplianceObjective::lastIterations)
        .def("updateCache", [](MultigridComplianceObjective &o, const Arr &x) {
                auto v = to_vec(x, o.tps->nel, "xPhys"); DevVec d; d.upload(v.data(), v.size()); o.updateCache(d);
            }, py::arg("xPhys"));

    // ------------------------------------------------------------------ problem, optimiser, LBL
    py::class_<Problem, PyProblem>(detail, "TopologyOptimizationProblem1_1_1")
        .def(py::init<TPSPtr, ObjectivePtr, std::vector<ConstraintPtr>, std::vector<FilterPtr>>())
        .def("evaluateObjective", &Problem::evaluateObjective)
        .def("evaluateObjectiveGradient", [](Problem &p) { return from_vec(p.evaluateObjectiveGradientAndReturn()); })
        .def("evaluateConstraints", [](Problem &p) { return from_vec(p.evaluateConstraints()); })
        .def("evaluateConstraintsJacobian", [](Problem &p) {
                std::vector<std::unique_ptr<DevVec>> rows;
                p.evaluateConstraintsJacobianDevice(rows);
                py::array_t<double> out({(py::ssize_t)rows.size(), (py::ssize_t)p.numVars()});
                for (size_t i = 0; i < rows.size(); ++i) rows[i]->download(out.mutable_data() + i * p.numVars());
                return out;
            })
        .def("numVars", &Problem::numVars)
        .def("getVars", [](Problem &p) { return from_vec(p.chain->design().host()); })
        .def("setVars", [](Problem &p, const Arr &x, bool force) {
                return p.setVars(to_vec(x, p.numVars(), "design variable"), force);
            }, py::arg("x"), py::arg("forceUpdate") = false)
        .def("getDensities", [](Problem &p) { return from_vec(p.tps->getDensities()); })
        .def_property("objective", [](Problem &p) { return p.objective; }, [](Problem &p, ObjectivePtr o) { p.objective = o; })
        .def_property_readonly("filters", [](Problem &p) { return p.filterList; })
        .def_property_readonly("filterChain", [](Problem &p) { return p.chain; })
        .def_property("constraints", [](Problem &p) { return p.constraints; },
                      [](Problem &p, std::vector<ConstraintPtr> c) { p.constraints = std::move(c); });

    py::class_<OCOptimizer>(detail, "OCOptimizer1_1_1")
        .def(py::init<Problem &>(), py::arg("problem"), py::keep_alive<1, 2>())
        .def("step", &OCOptimizer::step, py::arg("m") = 0.2, py::arg("p") = 0.5, py::arg("ctol") = 1e-6,
             py::arg("inplace") = true)
        .def_readonly("lastBisectionSteps", &OCOptimizer::lastBisectionSteps);

    py::class_<LBL>(detail, "LayerByLayerEvaluator1_1_1")
        .def("selectInitMethod", &LBL::selectInitMethod, py::arg("method"))
        .def("run", &LBL::run, py::arg("solver"), py::arg("zeroInit"), py::arg("layerIncrement"), py::arg("maxIter"),
             py::arg("tol"), py::arg("it_callback") = nullptr, py::arg("mgIterations") = 1,
             py::arg("mgSmoothingIterations") = 1, py::arg("fullMultigrid") = false, py::arg("verbose") = false,
             py::arg("lblCallback") = py::none())
        .def("objective", [](const LBL &l) { return l.total; })
        .def("gradient", [](const LBL &l) { return from_vec(l.grad); })
        .def("perLayerCompliance", [](const LBL &l) { return l.perLayer; })
        .def("perLayerIterations", [](const LBL &l) { return l.perLayerIters; })
        .def_property("precision", [](const LBL &l) { return l.precision == VF_PREC_FP32 ? "fp32" : "fp64"; },
                      [](LBL &l, const std::string &p) { l.precision = p == "fp32" ? VF_PREC_FP32 : VF_PREC_FP64; })
        .def_readwrite("batched", &LBL::batched)
        .def_readwrite("pipelines", &LBL::pipelines)
        .def_readwrite("useGraphs", &LBL::useGraphs)
        .def_readonly("lastRunSeconds", &LBL::seconds);

    // ------------------------------------------------------------------ factories (VoxelFEM.cc:221-298, 423-430)
    m.def("TensorProductSimulator", [](const std::vector<size_t> &degrees, const std::array<std::vector<double>, 2> &bbox,
                                       const std::vector<int64_t> &elems, NumberType nt) {
            if (degrees == std::vector<size_t>{1, 1} && nt == NumberType::DOUBLE && elems.size() == 2) {
                if (bbox[0].size() != 2 || bbox[1].size() != 2) throw std::runtime_error("domainBBox must hold two 2D corners");
                return makeTPS({bbox[0][0], bbox[0][1], 0.0}, {bbox[1][0], bbox[1][1], 1.0}, {elems[0], elems[1], 1}, 2);
            }
            if (degrees != std::vector<size_t>{1, 1, 1} || nt != NumberType::DOUBLE || elems.size() != 3)
                throw std::runtime_error("No template instantiation matching degreesPerDimension/number type!");
            if (bbox[0].size() != 3 || bbox[1].size() != 3) throw std::runtime_error("domainBBox must hold two 3D corners");
            return makeTPS({bbox[0][0], bbox[0][1], bbox[0][2]}, {bbox[1][0], bbox[1][1], bbox[1][2]},
                           {elems[0], elems[1], elems[2]});
        }, py::arg("degreesPerDimension"), py::arg("domainBBox"), py::arg("elementsPerDimension"),
        py::arg("numberType") = NumberType::DOUBLE);
    m.def("TopologyOptimizationProblem", [](TPSPtr tps, ObjectivePtr o, std::vector<ConstraintPtr> c, std::vector<FilterPtr> f) {
            return std::make_unique<Problem>(tps, o, c, f);
        }, py::arg("simulator"), py::arg("objective"), py::arg("constraints"), py::arg("filters"));
    m.def("ComplianceObjective", [](TPSPtr t) { return std::make_shared<ComplianceObjective>(t); }, py::arg("simulator"));
    m.def("MultigridComplianceObjective", [](MGPtr mg) { return std::make_shared<MultigridComplianceObjective>(mg); },
          py::arg("mg_solver"));
    m.def("LayerByLayerEvaluator", [](TPSPtr t) { return std::make_unique<LBL>(t); }, py::arg("lblSim"));
    m.def("OCOptimizer", [](Problem &p) { return std::make_unique<OCOptimizer>(p); }, py::keep_alive<0, 1>());
    m.def("getClassName", [](TPSPtr t, const std::string &name) {
            return "pyVoxelFEM.detail." + name + (t && t->N == 2 ? "1_1" : "1_1_1");
        }, py::arg("simulator"), py::arg("name"));
    // the planar instantiation's classes: the same Python types (one implementation serves both dimensions)
    for (const char *name : {"TensorProductSimulator", "MultigridSolver", "Objective", "ComplianceObjective",
                             "MultigridComplianceObjective", "TopologyOptimizationProblem", "OCOptimizer",
                             "LayerByLayerEvaluator"}) {
        const std::string n3 = std::string(name) + "1_1_1";
        if (py::hasattr(detail, n3.c_str())) detail.attr((std::string(name) + "1_1").c_str()) = detail.attr(n3.c_str());
    }
    m.def("backendVersion", []() { return std::string(vf_version()); });
}
