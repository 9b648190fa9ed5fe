// tortoise_host.hpp -- C++ host layer with the reference's Tortoise:: types,
// backed by the MI355X collision operator of libtortoise_gpu.so (C ABI in
// include/tortoise_gpu.h).
//
// Mirrors, for the path the GPU backend evaluates (one electron band on a
// structured box mesh, e-ph scattering with a dispersionless phonon bath):
//   Region<NDim>              include/Geometry/Structured/DirectSpace/Region.hpp:60-111
//   Mesh<NDim>                include/Geometry/Structured/DirectSpace/Mesh.hpp:73-230
//   Function<NDim>            include/Geometry/Structured/FunctionSpace/Function.hpp:67-262
//   Band<NDim>                include/Physics/Material/Band.hpp:49-113
//   ScatteringChannel<NDim,3> include/Physics/Scattering/ScatteringChannel.hpp:54-192
//   Material<NDim>            include/Physics/Material/Material.hpp:58-188
//   MaterialStatus<NDim>      include/Physics/Status/MaterialStatus.hpp:46-140
//   MaterialTimeStatus<NDim>  include/Physics/Status/MaterialTimeStatus.hpp:55-90
//
// Data model kept from the reference: a Function stores, per element, the
// LinearForm [f(v0), f(v1)-f(v0), ...] (element-major, Function.hpp:96,
// :326-344), vertex order and element numbering of Mesh.cpp:55-87,:384-433.
// The backend works on nodal values of the (nsec+1)^NDim vertex lattice; the
// DG -> nodal gather reads c0 (+ cj) of any element touching the node, which
// is exact for the continuous fields Function(mesh, lambda) builds
// (Function.cpp:60-68); nodal -> DG writes c0 = g(v0), cj = g(vj) - g(v0).
//
// Errors from the backend are thrown as std::runtime_error with the
// library's message.  Plotting is not provided (no gnuplot pipes).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <deque>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/tortoise_gpu.h"

namespace Tortoise {

using Real = double;
constexpr Real fermion = -1.0;   // Band statistics (Band.hpp:58)
constexpr Real boson = 1.0;
constexpr int in = 1;            // leg directions (Geometry.hpp:119)
constexpr int out = -1;

// ------------------------------------------------------------------- Point
template <int NDim>
struct Point {
    std::array<Real, NDim> c{};
    Point() = default;
    Point(std::initializer_list<Real> v) {
        int i = 0;
        for (Real x : v) if (i < NDim) c[i++] = x;
    }
    Real& operator()(int i) { return c[i]; }
    Real operator()(int i) const { return c[i]; }
    Real squaredNorm() const { Real s = 0; for (Real x : c) s += x * x; return s; }
    Real norm() const { return std::sqrt(squaredNorm()); }
    Point operator+(const Point& o) const { Point r; for (int i = 0; i < NDim; ++i) r.c[i] = c[i] + o.c[i]; return r; }
    Point operator-(const Point& o) const { Point r; for (int i = 0; i < NDim; ++i) r.c[i] = c[i] - o.c[i]; return r; }
    Point operator*(Real s) const { Point r; for (int i = 0; i < NDim; ++i) r.c[i] = c[i] * s; return r; }
};
template <int NDim>
Point<NDim> operator*(Real s, const Point<NDim>& p) { return p * s; }
template <int NDim>
Real dot(const Point<NDim>& a, const Point<NDim>& b) { Real s = 0; for (int i = 0; i < NDim; ++i) s += a(i) * b(i); return s; }

inline int factorial(int d) { int r = 1; for (int i = 2; i <= d; ++i) r *= i; return r; }

// ------------------------------------------------------------------ Region
// The parallelepiped origin + sum_a t_a gVec[a], t in [0,1]^NDim
// (Region.hpp:29-34): the zone whose sides the umklapp vectors combine.
template <int NDim>
class Region {
public:
    Point<NDim> origin;
    std::array<Point<NDim>, NDim> gVec;   // gVec[a]: side a (Region::gVec column a)
    Region(const Point<NDim>& t_origin, const std::array<Point<NDim>, NDim>& t_gVec)
        : origin(t_origin), gVec(t_gVec) {
        if (!(std::fabs(sideDeterminant(gVec)) > 0.0))
            throw std::invalid_argument("Region: the sides are degenerate");
    }
    static Real sideDeterminant(const std::array<Point<NDim>, NDim>& v) {
        if constexpr (NDim == 1) return v[0](0);
        else if constexpr (NDim == 2) return v[0](0) * v[1](1) - v[1](0) * v[0](1);
        else
            return v[0](0) * (v[1](1) * v[2](2) - v[2](1) * v[1](2)) -
                   v[1](0) * (v[0](1) * v[2](2) - v[2](1) * v[0](2)) +
                   v[2](0) * (v[0](1) * v[1](2) - v[1](1) * v[0](2));
    }
    bool axisAligned() const {
        for (int a = 0; a < NDim; ++a)
            for (int d = 0; d < NDim; ++d)
                if (a != d && gVec[a](d) != 0.0) return false;
        return true;
    }
};

// -------------------------------------------------------------------- Mesh
template <int NDim>
class Mesh {
public:
    static constexpr int nVertElement = NDim + 1;
    const int nElementSection = factorial(NDim);
    const Region<NDim>* region;
    Point<NDim> origin;          // physical origin of the mesh
    std::array<Point<NDim>, NDim> sides;   // its side vectors (Mesh::gVec columns)
    Point<NDim> extent;          // axis-aligned mesh: the sides' lengths along the axes
    bool axisAligned = true;
    std::array<int, NDim> nSecSplits;
    int numberSections, numberElements, dataVectorDim;
    Real elemVolume;

    // Mesh.cpp:268-293: origin = region.origin + gVec * relOrigin, the sides
    // gVec * diag(relSize) (terms in axis order: on an axis-aligned region
    // the other sides add exact zeros)
    Mesh(const Region<NDim>& t_region, const Point<NDim>& relOrigin, const Point<NDim>& relSize,
         const std::array<int, NDim>& t_nSecSplits)
        : region(&t_region), nSecSplits(t_nSecSplits) {
        numberSections = 1;
        for (int d = 0; d < NDim; ++d) {
            Real o = t_region.origin(d);
            for (int a = 0; a < NDim; ++a) o = o + t_region.gVec[a](d) * relOrigin(a);
            origin(d) = o;
            numberSections *= nSecSplits[d];
        }
        for (int a = 0; a < NDim; ++a) sides[a] = t_region.gVec[a] * relSize(a);
        axisAligned = t_region.axisAligned();
        Real vol = 1.0;
        if (axisAligned) {
            for (int d = 0; d < NDim; ++d) {
                extent(d) = sides[d](d);
                vol *= extent(d) / nSecSplits[d];
            }
        } else {
            vol = Region<NDim>::sideDeterminant(sides);
            for (int d = 0; d < NDim; ++d) {
                extent(d) = 0.0;
                vol /= nSecSplits[d];
            }
        }
        numberElements = numberSections * nElementSection;
        dataVectorDim = numberElements * nVertElement;
        elemVolume = std::fabs(vol / nElementSection);
    }
    Mesh(const Mesh&) = delete;
    Mesh(Region<NDim>&&, const Point<NDim>&, const Point<NDim>&, const std::array<int, NDim>&) = delete;

    int64_t numberNodes() const {
        int64_t n = 1;
        for (int d = 0; d < NDim; ++d) n *= nSecSplits[d] + 1;
        return n;
    }
    // origin + sum_a (sides[a] / nSecSplits[a]) i_a, terms in axis order
    Point<NDim> node(int64_t id) const {
        std::array<int64_t, NDim> idx;
        for (int a = 0; a < NDim; ++a) {
            idx[a] = id % (nSecSplits[a] + 1);
            id /= (nSecSplits[a] + 1);
        }
        Point<NDim> p;
        for (int d = 0; d < NDim; ++d) {
            Real x = origin(d);
            for (int a = 0; a < NDim; ++a) x = x + (sides[a](d) / nSecSplits[a]) * (Real)idx[a];
            p(d) = x;
        }
        return p;
    }
    // node id of vertex `vert` of element `elem` (Mesh.cpp:55-87, :412-420)
    int64_t elementNode(int elem, int vert) const {
        int s = elem / nElementSection, t = elem % nElementSection;
        int sc[3] = {0, 0, 0};
        for (int d = 0; d < NDim; ++d) { sc[d] = s % nSecSplits[d]; s /= nSecSplits[d]; }
        int avail[3] = {0, 1, 2}, navail = NDim, rem = t, mask = 0;
        for (int k = 0; k < vert; ++k) {
            int f = factorial(NDim - 1 - k), idx = rem / f;
            rem %= f;
            mask |= 1 << avail[idx];
            for (int q = idx; q < navail - 1; ++q) avail[q] = avail[q + 1];
            --navail;
        }
        int64_t id = 0, stride = 1;
        for (int d = 0; d < NDim; ++d) {
            id += stride * (sc[d] + ((mask >> d) & 1));
            stride *= nSecSplits[d] + 1;
        }
        return id;
    }
    Point<NDim> operator()(int elem, int vert) const { return node(elementNode(elem, vert)); }
    Point<NDim> elemCentre(int elem) const {
        Point<NDim> p;
        for (int v = 0; v < nVertElement; ++v) p = p + (*this)(elem, v);
        return p * (1.0 / nVertElement);
    }
};

// ---------------------------------------------------------------- Function
template <int NDim>
class Function {
public:
    const Mesh<NDim>* mesh;
    std::vector<Real> vec;   // per element [f(v0), f(v1)-f(v0), ...]

    explicit Function(const Mesh<NDim>& m) : mesh(&m), vec(m.dataVectorDim, 0.0) {}
    Function(const Mesh<NDim>& m, Real value) : mesh(&m), vec(m.dataVectorDim, 0.0) {
        for (int e = 0; e < m.numberElements; ++e) vec[e * m.nVertElement] = value;
    }
    Function(const Mesh<NDim>& m, const std::function<Real(Point<NDim>)>& f) : mesh(&m), vec(m.dataVectorDim) {
        for (int e = 0; e < m.numberElements; ++e)
            for (int v = 0; v < m.nVertElement; ++v) vec[e * m.nVertElement + v] = f(m(e, v));
        toLinFormFromNodal();
    }
    static Function fromNodal(const Mesh<NDim>& m, const std::vector<Real>& nodal) {
        Function r(m);
        for (int e = 0; e < m.numberElements; ++e)
            for (int v = 0; v < m.nVertElement; ++v)
                r.vec[e * m.nVertElement + v] = nodal[m.elementNode(e, v)];
        r.toLinFormFromNodal();
        return r;
    }
    // nodal values on the vertex lattice (exact for continuous fields)
    std::vector<Real> toNodal() const {
        std::vector<Real> out(mesh->numberNodes(), 0.0);
        const int nv = mesh->nVertElement;
        for (int e = 0; e < mesh->numberElements; ++e)
            for (int v = 0; v < nv; ++v)
                out[mesh->elementNode(e, v)] = vec[e * nv] + (v ? vec[e * nv + v] : 0.0);
        return out;
    }
    void toLinFormFromNodal() {
        const int nv = mesh->nVertElement;
        for (int e = 0; e < mesh->numberElements; ++e)
            for (int v = 1; v < nv; ++v) vec[e * nv + v] -= vec[e * nv];
    }
    void toNodalFromLinForm() {
        const int nv = mesh->nVertElement;
        for (int e = 0; e < mesh->numberElements; ++e)
            for (int v = 1; v < nv; ++v) vec[e * nv + v] += vec[e * nv];
    }
    // Function.cpp:123-234: linear ops on coefficients, nonlinear ops nodal
    Function& operator+=(const Function& o) { for (size_t i = 0; i < vec.size(); ++i) vec[i] += o.vec[i]; return *this; }
    Function& operator-=(const Function& o) { for (size_t i = 0; i < vec.size(); ++i) vec[i] -= o.vec[i]; return *this; }
    Function& operator*=(Real s) { for (auto& x : vec) x *= s; return *this; }
    Function& operator+=(Real s) { for (int e = 0; e < mesh->numberElements; ++e) vec[e * mesh->nVertElement] += s; return *this; }
    Function& operator*=(const Function& o) { return nodalOp(o, [](Real a, Real b) { return a * b; }); }
    Function& operator/=(const Function& o) { return nodalOp(o, [](Real a, Real b) { return a / b; }); }
    Function operator+(const Function& o) const { Function r(*this); r += o; return r; }
    Function operator-(const Function& o) const { Function r(*this); r -= o; return r; }
    Function operator*(const Function& o) const { Function r(*this); r *= o; return r; }
    Function operator*(Real s) const { Function r(*this); r *= s; return r; }
    friend Function operator*(Real s, const Function& f) { return f * s; }
    Function map(const std::function<Real(Real)>& g) const {
        Function r(*this);
        r.toNodalFromLinForm();
        for (auto& x : r.vec) x = g(x);
        r.toLinFormFromNodal();
        return r;
    }
    // Function.cpp:369-372
    Real integrate() const {
        Real s = 0, s0 = 0;
        const int nv = mesh->nVertElement;
        for (int e = 0; e < mesh->numberElements; ++e) {
            s0 += vec[e * nv];
            for (int v = 0; v < nv; ++v) s += vec[e * nv + v];
        }
        return mesh->elemVolume * (s + NDim * s0) / (NDim + 1);
    }
    // Function.cpp:373-375 (order-2 Gauss is exact for P1*P1)
    Real integrate(const Function& rhs) const {
        const int nv = mesh->nVertElement;
        Real tot = 0;
        for (int e = 0; e < mesh->numberElements; ++e) {
            Real a[4], b[4], sa = 0, sb = 0, sab = 0;
            for (int v = 0; v < nv; ++v) {
                a[v] = vec[e * nv] + (v ? vec[e * nv + v] : 0.0);
                b[v] = rhs.vec[e * nv] + (v ? rhs.vec[e * nv + v] : 0.0);
                sa += a[v];
                sb += b[v];
                sab += a[v] * b[v];
            }
            tot += sab + sa * sb;
        }
        return mesh->elemVolume * tot / ((NDim + 1) * (NDim + 2));
    }

private:
    Function& nodalOp(const Function& o, Real (*op)(Real, Real)) {
        toNodalFromLinForm();
        Function b(o);
        b.toNodalFromLinForm();
        for (size_t i = 0; i < vec.size(); ++i) vec[i] = op(vec[i], b.vec[i]);
        toLinFormFromNodal();
        return *this;
    }
};

// -------------------------------------------------------------------- Band
template <int NDim>
class Band {
public:
    std::string name;
    Real charge, statistics;
    const Mesh<NDim>* mesh;
    Function<NDim> dispersion;
    bool constrained = false;
    std::function<Function<NDim>(Real)> constrainedPopulation;
    Real absoluteErrorTolerance = 1.e-3;         // Band.hpp:77-82
    Real relativeErrorTolerance = 1.e-3;
    Real penalisationUnphysicalPopulation = 0.;
    // The device holds a constrained (uniform) population as samples of
    // constrainedPopulation(t), linear in between: samples per dense step of
    // a propagation (a backend parameter; the reference calls the function
    // itself at every stage time, MaterialStatus.cpp:198-206).
    int constraintSamplesPerDenseStep = 64;
    Band(std::string n, Real ch, Real st, const Mesh<NDim>& m, const Function<NDim>& disp)
        : name(std::move(n)), charge(ch), statistics(st), mesh(&m), dispersion(disp) {}
};

// ------------------------------------------------------ ScatteringChannel
// Two channels the backend evaluates:
//   legs == 3: e-ph {electron in, phonon in, electron out} with the phonon
//     band acting as a bath of a dispersionless mode (the W-form);
//   legs == 4: e-e {el in, el in, el out, el out} with the screened Coulomb
//     amplitude V(q)^2, V(q) = 1/(|q|^2 + screening), on the periodic lattice
//     of the zone (the functAmplitude of a 4-leg channel,
//     ScatteringChannel.cpp:399-415).
// `amplitude` is g^2 (ScatteringChannel.hpp:70); the broadening and the
// energy grid of the delta are new parameters.
template <int NDim>
class ScatteringChannel {
public:
    int legs = 3;
    std::string electronBand, phononBand;
    Real screening = 0.09;          // e-e: kappa^2 of V(q)
    Real amplitude = 1.0;
    Real deltaBroadening = 0.01;
    Real energyGridStep = 0.0;      // <= 0: deltaBroadening / 8
    Real deltaCutoff = 8.0;         // in units of deltaBroadening
    // the reference operator (Material::referenceOperator): MC points per
    // tuple, screening thresholds and umklapp vectors {-n..n}^NDim
    // (Material.hpp:89-91, ScatteringChannel.hpp:80-83)
    int numMCPoints = 50;
    Real minimumDeterminant = 1e-9;
    Real scattElemeThreshold = 1e-9;
    int umklappSteps = 0;
};

template <int NDim> class MaterialStatus;

// ---------------------------------------------------------------- Material
template <int NDim>
class Material {
public:
    std::string name;
    const Region<NDim>* region;
    std::deque<Band<NDim>> bands;                       // insertion order (StringMap.hpp:78-79)
    std::deque<ScatteringChannel<NDim>> channels;

    // false: the W-form (nodal P1, broadened delta, assembled once);
    // true: the reference's own operator -- element tuples and the Monte-Carlo
    // delta-shell integral per tuple on the DG coefficients (tg_tuple_*)
    bool referenceOperator = false;
    // with referenceOperator: the phonon band evolves (its derivative is
    // returned and integrated) instead of acting as a bath
    bool dynamicPhonons = false;
    uint64_t mcSeed = 0;     // start of the MC point stream (the reference uses std::rand)
    // W-form occupation factors of the test node: true, the hat average of the
    // P1 occupation (its energy relaxation follows the reference operator's,
    // which integrates those P1 products exactly), with the number defect
    // removed; false, the nodal value (the lumped form)
    bool hatOccupation = true;
    int32_t formFlags() const { return hatOccupation ? TG_CONSERVE : 0; }

    Material(std::string n, const Region<NDim>& r) : name(std::move(n)), region(&r) {}
    Material(const Material&) = delete;
    ~Material() {
        if (ctx_) tg_ctx_destroy(ctx_);
        if (tctx_) tg_tuple_destroy(tctx_);
    }

    void addBand(const std::string& n, Real charge, Real stat, const Mesh<NDim>& m,
                 const std::function<Real(Point<NDim>)>& disp) {
        bands.emplace_back(n, charge, stat, m, Function<NDim>(m, disp));
        invalidate();
    }
    void addBand(const std::string& n, Real charge, Real stat, const Mesh<NDim>& m, Real disp) {
        bands.emplace_back(n, charge, stat, m, Function<NDim>(m, disp));
        invalidate();
    }
    Band<NDim>& operator[](const std::string& n) {
        for (auto& b : bands) if (b.name == n) return b;
        throw std::out_of_range("no band " + n);
    }
    const Band<NDim>& operator[](const std::string& n) const {
        for (auto& b : bands) if (b.name == n) return b;
        throw std::out_of_range("no band " + n);
    }
    // wildcard match (StringUtilities.cpp:44-59): '*' matches any suffix
    std::vector<std::string> matchListId(const std::string& pat) const {
        std::vector<std::string> r;
        for (auto& b : bands) if (match(pat, b.name)) r.push_back(b.name);
        return r;
    }
    // Material.cpp:244-301 for {el:*, pn:*, el:*} with {in, in, out} (e-ph)
    // and {el:*, el:*, el:*, el:*} with {in, in, out, out} (e-e, all four legs
    // on one fermion band)
    template <int NLegs>
    std::vector<ScatteringChannel<NDim>*> addScatteringChannelsWildcards(
        const std::array<std::string, NLegs>& legs, const std::array<int, NLegs>& dirs, int nUmklapp) {
        static_assert(NLegs == 3 || NLegs == 4, "the backend evaluates 3-leg e-ph and 4-leg e-e channels");
        std::vector<ScatteringChannel<NDim>*> r;
        if constexpr (NLegs == 3) {
            if (!(dirs[0] == in && dirs[1] == in && dirs[2] == out))
                throw std::invalid_argument("e-ph channel must be {in, in, out}");
            for (auto& e : matchListId(legs[0]))
                for (auto& p : matchListId(legs[1])) {
                    if ((*this)[e].statistics != fermion || (*this)[p].statistics != boson) continue;
                    channels.push_back(ScatteringChannel<NDim>{});
                    channels.back().electronBand = e;
                    channels.back().phononBand = p;
                    channels.back().umklappSteps = nUmklapp;
                    r.push_back(&channels.back());
                }
        } else {
            if (!(dirs[0] == in && dirs[1] == in && dirs[2] == out && dirs[3] == out))
                throw std::invalid_argument("e-e channel must be {in, in, out, out}");
            for (auto& e : matchListId(legs[0])) {
                if ((*this)[e].statistics != fermion) continue;
                bool all = true;
                for (int l = 1; l < 4; ++l) {
                    const auto ids = matchListId(legs[l]);
                    all = all && std::find(ids.begin(), ids.end(), e) != ids.end();
                }
                if (!all) continue;
                channels.push_back(ScatteringChannel<NDim>{});
                channels.back().legs = 4;
                channels.back().electronBand = e;
                channels.back().deltaBroadening = 0.02;
                r.push_back(&channels.back());
            }
        }
        invalidate();
        return r;
    }

    // ---- backend
    // Context of the e-ph channel (exactly one) and the e-e channel (at most
    // one, on the same electron band); built on first use with the channel
    // parameters current at that time.
    tg_ctx* backend(Real bath_n, int max_batch = 1) {
        if (ctx_ && ctx_bath_ == bath_n) return ctx_;
        const ScatteringChannel<NDim>* eph = nullptr;
        const ScatteringChannel<NDim>* ee = nullptr;
        for (const auto& c : channels) {
            if (c.legs == 3) {
                if (eph) throw std::runtime_error("the backend evaluates one e-ph channel");
                eph = &c;
            } else {
                if (ee) throw std::runtime_error("the backend evaluates at most one e-e channel");
                ee = &c;
            }
        }
        if (!eph) throw std::runtime_error("the backend needs an e-ph channel");
        if (ee && ee->electronBand != eph->electronBand)
            throw std::runtime_error("the e-e channel must act on the e-ph channel's electron band");
        const auto& ch = *eph;
        const Band<NDim>& el = (*this)[ch.electronBand];
        const Band<NDim>& pn = (*this)[ch.phononBand];
        const auto omega = pn.dispersion.toNodal();
        for (Real w : omega)
            if (std::fabs(w - omega[0]) > 1e-12 * std::fabs(omega[0]))
                throw std::runtime_error("phonon band must be dispersionless");
        if (!el.mesh->axisAligned)
            throw std::runtime_error("the W-form backend meshes axis-aligned zones; the reference operator "
                                     "(tupleBackend) takes general ones");
        eps_ = el.dispersion.toNodal();
        tg_problem_desc d{};
        d.dim = NDim;
        for (int q = 0; q < 3; ++q) {
            d.nsec[q] = q < NDim ? el.mesh->nSecSplits[q] : 1;
            d.origin[q] = q < NDim ? el.mesh->origin(q) : 0.0;
            d.extent[q] = q < NDim ? el.mesh->extent(q) : 1.0;
        }
        d.eps = eps_.data();
        d.g = std::sqrt(ch.amplitude);
        d.omega0 = omega[0];
        d.sigma = ch.deltaBroadening;
        d.bath_n = bath_n;
        d.de = ch.energyGridStep;
        d.cutoff = ch.deltaCutoff;
        d.max_batch = max_batch;
        d.occupation = hatOccupation ? TG_OCC_HAT : TG_OCC_NODAL;
        if (ee) {
            d.ee_g = std::sqrt(ee->amplitude);
            d.ee_kappa2 = ee->screening;
            d.ee_sigma = ee->deltaBroadening;
            d.ee_dw = ee->energyGridStep;
        }
        if (ctx_) tg_ctx_destroy(ctx_);
        ctx_ = nullptr;
        check(tg_ctx_create(&d, &ctx_));
        check(tg_kernel_assemble(ctx_, 0));
        ctx_bath_ = bath_n;
        return ctx_;
    }
    void check(tg_status s) const {
        if (s != TG_OK) throw std::runtime_error(std::string("tortoise_gpu: ") + tg_last_error(ctx_));
    }
    void checkTuple(tg_status s) const {
        if (s != TG_OK) throw std::runtime_error(std::string("tortoise_gpu: ") + tg_tuple_last_error(tctx_));
    }

    // The reference operator of the e-ph channel: tuple list built on first
    // use (ScatteringChannel::precalculateIndices, ScatteringChannel.cpp:421-538)
    // with the MC points of the first constructMCPoints().
    tg_tuple_ctx* tupleBackend() {
        if (tctx_) return tctx_;
        const ScatteringChannel<NDim>* eph = nullptr;
        for (const auto& c : channels) {
            if (c.legs == 4)
                throw std::runtime_error(
                    "the reference operator evaluates one e-ph channel; a 4-leg channel in tuple form is "
                    "tg_tuple4_create, which this layer does not combine with the e-ph channel");
            if (eph) throw std::runtime_error("the reference operator evaluates one e-ph channel");
            eph = &c;
        }
        if (!eph) throw std::runtime_error("the reference operator needs an e-ph channel");
        const Band<NDim>& el = (*this)[eph->electronBand];
        omegaPh_ = (*this)[eph->phononBand].dispersion.toNodal();
        bool flat = true;
        for (Real w : omegaPh_) flat = flat && w == omegaPh_[0];
        eps_ = el.dispersion.toNodal();
        tg_tuple_desc d{};
        d.dim = NDim;
        for (int q = 0; q < 3; ++q) {
            d.nsec[q] = q < NDim ? el.mesh->nSecSplits[q] : 1;
            d.origin[q] = q < NDim ? el.mesh->origin(q) : 0.0;
            d.extent[q] = q < NDim ? el.mesh->extent(q) : 1.0;
        }
        // the mesh's sides and the region's (the umklapp basis,
        // ScatteringChannel.cpp:330-334): a sub-mesh of a region moves by the
        // region's sides, not its own
        for (int a = 0; a < NDim; ++a)
            for (int q = 0; q < NDim; ++q) {
                d.sides[a][q] = el.mesh->sides[a](q);
                d.zone[a][q] = el.mesh->region->gVec[a](q);
            }
        d.eps = eps_.data();
        d.omega0 = omegaPh_[0];
        d.eps_ph = flat ? nullptr : omegaPh_.data();   // a dispersive phonon band
        d.dynamic_phonons = dynamicPhonons ? 1 : 0;
        d.amplitude = eph->amplitude;
        d.n_umklapp = eph->umklappSteps;
        d.num_mc = eph->numMCPoints;
        d.min_det = eph->minimumDeterminant;
        d.threshold = eph->scattElemeThreshold;
        d.max_batch = 1;
        tg_status s = tg_tuple_create(&d, &tctx_);
        if (s != TG_OK) {
            std::string msg = tctx_ ? tg_tuple_last_error(tctx_) : "tg_tuple_create failed";
            if (tctx_) tg_tuple_destroy(tctx_);
            tctx_ = nullptr;
            throw std::runtime_error("tortoise_gpu: " + msg);
        }
        tupleChannel_ = eph;
        return tctx_;
    }
    // Material::constructMCPoints: new points for the e-ph channel (leg 0 in
    // the reference simplex, the partially free leg in the (NDim-1)-simplex,
    // ScatteringChannel.cpp:677-688, Point.hpp:44-68), from a splitmix64 stream.
    void constructMCPoints() {
        tg_tuple_ctx* t = tupleBackend();
        const int nmc = tupleChannel_->numMCPoints;
        const int rows = tg_tuple_point_rows(NDim, 3);
        std::vector<Real> pts((size_t)rows * nmc, 0.0);
        for (int m = 0; m < nmc; ++m) {
            Real x[3];
            simplexPoint(NDim, x);
            for (int d = 0; d < NDim; ++d) pts[(size_t)d * nmc + m] = x[d];
        }
        if (NDim > 1)
            for (int m = 0; m < nmc; ++m) {
                Real x[3];
                simplexPoint(NDim - 1, x);
                for (int d = 0; d < NDim - 1; ++d) pts[(size_t)(rows - (NDim - 1) + d) * nmc + m] = x[d];
            }
        checkTuple(tg_tuple_set_points(t, pts.data(), 0));
        havePoints_ = true;
    }
    bool haveMCPoints() const { return havePoints_; }
    const ScatteringChannel<NDim>& tupleChannel() const { return *tupleChannel_; }

    // Material.cpp:396-402 -> ScatteringChannel::scatteringRates(status, band)
    // (ScatteringChannel.cpp:544-612): the e-ph channel's rates of `bandName`
    // (the electron band) from the reference operator, summed over the legs
    // the band occupies.  Channel patterns select the e-ph channel.
    template <int NLegs>
    Function<NDim> scatteringRates(const std::array<std::string, NLegs>& bandsNamesWildCards,
                                   const std::array<int, NLegs>& legDirection,
                                   const MaterialStatus<NDim>& status, const std::string& bandName);

private:
    tg_ctx* ctx_ = nullptr;
    Real ctx_bath_ = -1.0;
    std::vector<Real> eps_;
    tg_tuple_ctx* tctx_ = nullptr;
    std::vector<Real> omegaPh_;
    const ScatteringChannel<NDim>* tupleChannel_ = nullptr;
    bool havePoints_ = false;
    uint64_t rngState_ = 0;
    bool rngStarted_ = false;
    void invalidate() {
        if (ctx_) tg_ctx_destroy(ctx_);
        ctx_ = nullptr;
        if (tctx_) tg_tuple_destroy(tctx_);
        tctx_ = nullptr;
        havePoints_ = false;
    }
    Real unitDraw() {
        if (!rngStarted_) { rngState_ = mcSeed; rngStarted_ = true; }
        uint64_t z = (rngState_ += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (Real)(z >> 11) * 0x1.0p-53;
    }
    void simplexPoint(int d, Real* x) {
        for (int i = 0; i < d; ++i) x[i] = unitDraw();
        if (d >= 2 && x[0] + x[1] > 1.0) { x[0] = 1.0 - x[0]; x[1] = 1.0 - x[1]; }
        if (d == 3) {
            if (x[1] + x[2] > 1.0) {
                Real t = x[2];
                x[2] = 1.0 - x[0] - x[1];
                x[1] = 1.0 - t;
            } else if (x[0] + x[1] + x[2] > 1.0) {
                Real t = x[2];
                x[2] = x[0] + x[1] + x[2] - 1.0;
                x[0] = 1.0 - x[1] - t;
            }
        }
    }
    static bool match(const std::string& pat, const std::string& s) {
        auto star = pat.find('*');
        if (star == std::string::npos) return pat == s;
        return s.compare(0, star, pat, 0, star) == 0;
    }
};

// ---------------------------------------------------------- MaterialStatus
template <int NDim>
class MaterialStatus {
public:
    Material<NDim>* const material;
    std::map<std::string, Function<NDim>> pop;

    explicit MaterialStatus(Material<NDim>& m) : material(&m) {
        for (auto& b : m.bands) pop.emplace(b.name, Function<NDim>(*b.mesh));
    }
    MaterialStatus(const MaterialStatus&) = default;
    MaterialStatus& operator=(const MaterialStatus& o) { pop = o.pop; return *this; }
    Function<NDim>& operator[](const std::string& n) { return pop.at(n); }
    const Function<NDim>& operator[](const std::string& n) const { return pop.at(n); }

    MaterialStatus& operator+=(const MaterialStatus& o) { for (auto& [k, f] : pop) f += o[k]; return *this; }
    MaterialStatus& operator*=(Real s) { for (auto& [k, f] : pop) f *= s; return *this; }

    // MaterialStatus.cpp:212-220
    void setToEquilibrium(Real T, Real muF, Real muB) {
        for (auto& b : material->bands) {
            const Real mu = b.statistics == fermion ? muF : muB;
            const Real s = b.statistics == fermion ? 1.0 : -1.0;
            pop.at(b.name) = b.dispersion.map([&](Real e) { return 1.0 / (std::exp((e - mu) / T) + s); });
        }
    }

    // MaterialStatus.cpp:146-161: d(populations)/dt of every channel.  The
    // phonon bath is constrained, its derivative is zero.
    MaterialStatus propagate() const {
        MaterialStatus r(*material);
        if (material->referenceOperator) {
            // ScatteringChannel::propagate (ScatteringChannel.cpp:651-670): the
            // tuple operator on the DG coefficients, inverse mass included
            tg_tuple_ctx* t = material->tupleBackend();
            if (!material->haveMCPoints()) material->constructMCPoints();
            const auto& tch = material->tupleChannel();
            const Function<NDim>& fe = (*this)[tch.electronBand];
            material->checkTuple(tg_tuple_set_amplitude(t, tch.amplitude));
            if (material->dynamicPhonons) {
                // both bands: [electrons | phonons]
                const Function<NDim>& fp = (*this)[tch.phononBand];
                std::vector<Real> y(fe.vec), o(2 * fe.vec.size());
                y.insert(y.end(), fp.vec.begin(), fp.vec.end());
                material->checkTuple(tg_tuple_propagate(t, y.data(), nullptr, o.data(),
                                                        fe.mesh->numberElements, 1, 0));
                Function<NDim> oe(*fe.mesh), op(*fp.mesh);
                std::copy(o.begin(), o.begin() + fe.vec.size(), oe.vec.begin());
                std::copy(o.begin() + fe.vec.size(), o.end(), op.vec.begin());
                r[tch.electronBand] = oe;
                r[tch.phononBand] = op;
                return r;
            }
            Function<NDim> out(*fe.mesh);
            const Real n = bathOccupation();
            material->checkTuple(tg_tuple_propagate(t, fe.vec.data(), &n, out.vec.data(),
                                                    fe.mesh->numberElements, 1, 0));
            r[tch.electronBand] = out;
            return r;
        }
        const auto& ch = material->channels.front();
        tg_ctx* ctx = material->backend(bathOccupation());
        const Function<NDim>& fe = (*this)[ch.electronBand];
        std::vector<Real> f = fe.toNodal(), I(f.size());
        material->check(tg_collision_apply(ctx, f.data(), I.data(), (int64_t)f.size(), 1, nullptr,
                                           material->formFlags()));
        r[ch.electronBand] = Function<NDim>::fromNodal(*fe.mesh, I);
        return r;
    }
    // MaterialStatus.cpp:189-196 with Band::relativeError (Band.cpp:87-101)
    Real relativeError(const MaterialStatus& population) const {
        const auto& ch = material->channels.front();
        tg_ctx* ctx = material->backend(population.bathOccupation());
        const Band<NDim>& b = (*material)[ch.electronBand];
        std::vector<Real> e = (*this)[ch.electronBand].toNodal(), p = population[ch.electronBand].toNodal();
        Real out = 0;
        material->check(tg_relative_error(ctx, e.data(), p.data(), (int64_t)e.size(), 1,
                                          b.relativeErrorTolerance, b.absoluteErrorTolerance, &out, 0));
        return out;
    }
    // occupation of the phonon bath: the population of the phonon band, which
    // the backend's bath forms need uniform (a non-uniform one is refused,
    // not replaced by one of its values)
    Real bathOccupation() const {
        const auto& ch = material->channels.front();
        return uniformValue((*this)[ch.phononBand], "phonon band population");
    }
    // the value of a uniform population; throws when its nodal values differ
    static Real uniformValue(const Function<NDim>& f, const char* what) {
        const auto v = f.toNodal();
        if (v.empty()) return 0.0;
        for (Real x : v)
            if (!(std::fabs(x - v[0]) <= 1e-12 * std::fabs(v[0])))
                throw std::runtime_error(std::string("tortoise_gpu: the ") + what +
                                         " is not uniform; the backend's phonon bath takes one occupation");
        return v[0];
    }
};

template <int NDim>
template <int NLegs>
Function<NDim> Material<NDim>::scatteringRates(const std::array<std::string, NLegs>& bandsNamesWildCards,
                                               const std::array<int, NLegs>& legDirection,
                                               const MaterialStatus<NDim>& status, const std::string& bandName) {
    static_assert(NLegs == 3, "rates are evaluated for the 3-leg e-ph channel");
    if (!(legDirection[0] == in && legDirection[1] == in && legDirection[2] == out))
        throw std::invalid_argument("e-ph channel must be {in, in, out}");
    tg_tuple_ctx* t = tupleBackend();
    const auto& ch = tupleChannel();
    bool sel = false;
    for (auto& e : matchListId(bandsNamesWildCards[0]))
        for (auto& p : matchListId(bandsNamesWildCards[1]))
            sel = sel || (e == ch.electronBand && p == ch.phononBand);
    const Function<NDim>& fe = status[ch.electronBand];
    Function<NDim> out(*fe.mesh);
    if (!sel || bandName != ch.electronBand) return out;   // no selected channel acts on the band
    if (!haveMCPoints()) constructMCPoints();
    checkTuple(tg_tuple_set_amplitude(t, ch.amplitude));
    if (dynamicPhonons) {
        // the state holds both bands, [electrons | phonons]
        const Function<NDim>& fp = status[ch.phononBand];
        std::vector<Real> y(fe.vec);
        y.insert(y.end(), fp.vec.begin(), fp.vec.end());
        checkTuple(tg_tuple_rates(t, y.data(), nullptr, out.vec.data(), fe.mesh->numberElements, 1, 0));
        return out;
    }
    const Real n = status.bathOccupation();
    checkTuple(tg_tuple_rates(t, fe.vec.data(), &n, out.vec.data(), fe.mesh->numberElements, 1, 0));
    return out;
}

// ------------------------------------------------------ MaterialTimeStatus
template <int NDim>
class MaterialTimeStatus {
public:
    Material<NDim>* const material;
    std::vector<Real> argVec;
    std::vector<MaterialStatus<NDim>> valVec;
    bool quiet = false;
    tg_step_stats lastStats{};

    explicit MaterialTimeStatus(const MaterialStatus<NDim>& s) : material(s.material) {
        argVec.push_back(0.0);
        valVec.push_back(s);
    }
    int size() const { return (int)argVec.size(); }
    Real arg(int i) const { return argVec[i]; }
    Real argback() const { return argVec.back(); }
    MaterialStatus<NDim>& operator[](int i) { return valVec[i]; }
    MaterialStatus<NDim>& back() { return valVec.back(); }

    // MaterialTimeStatus.cpp:102-137 -> DormandPrince54.hpp:479-665: adaptive
    // DP5(4) from the last entry to finalTime; every dense output is appended.
    // A constrained phonon band (Band::constrained, a uniform population
    // n(t)) is the bath the device evaluates at every stage time, and the
    // step vetoer reads it (MaterialStatus.cpp:198-206, MaterialTimeStatus.cpp:60-100).
    void propagate(Real finalTime, Real denseStep) {
        if (material->referenceOperator) {
            propagateReference(finalTime, denseStep);
            return;
        }
        const MaterialStatus<NDim> start = valVec.back();
        const auto& ch = material->channels.front();
        const Band<NDim>& el = (*material)[ch.electronBand];
        const Band<NDim>& pb = (*material)[ch.phononBand];
        checkConstraints(el, pb);
        const Real t0 = argVec.back();
        tg_ctx* ctx = material->backend(pb.constrained ? constrainedValue(pb, t0) : start.bathOccupation());
        int32_t flags = (quiet ? TG_QUIET : 0) | material->formFlags();
        if (pb.constrained) {
            const Schedule sc = sampleConstraint(pb, t0, finalTime, denseStep);
            tg_bath_schedule d = sc.desc(pb);
            material->check(tg_set_bath_schedule(ctx, &d));
            flags |= TG_BATH_SCHEDULE;
        }
        std::vector<Real> f = start[ch.electronBand].toNodal();
        const int64_t n = (int64_t)f.size();
        const int maxDense = (int)std::floor((finalTime - t0) / denseStep + 1e-9) + 1;
        std::vector<Real> dense((size_t)maxDense * n);
        tg_status s = tg_dp5_integrate(ctx, f.data(), n, 1, t0, finalTime, denseStep,
                                       el.relativeErrorTolerance, el.absoluteErrorTolerance, nullptr,
                                       flags, dense.data(), maxDense, &lastStats, nullptr, 0);
        if (s != TG_OK && s != TG_ERR_STEP_UNDERFLOW) material->check(s);
        if (s == TG_ERR_STEP_UNDERFLOW) std::printf("The step size has become too small to be handled\n");
        const int nd = lastStats.n_dense < maxDense ? lastStats.n_dense : maxDense;
        for (int k = 0; k < nd; ++k) {
            MaterialStatus<NDim> st = start;
            st[ch.electronBand] = Function<NDim>::fromNodal(
                *el.mesh, std::vector<Real>(dense.begin() + (size_t)k * n, dense.begin() + (size_t)(k + 1) * n));
            const Real td = t0 + (k + 1) * denseStep;
            if (pb.constrained) st[ch.phononBand] = pb.constrainedPopulation(td);   // enforceValue at td
            argVec.push_back(td);
            valVec.push_back(std::move(st));
        }
    }
    // the name the reference's examples call (examples/main.cpp:99)
    void propagateDeterministic(Real finalTime, Real denseStep) { propagate(finalTime, denseStep); }

private:
    // samples of a constrained uniform population from t0 over twice the
    // propagation span (the last step may end past finalTime; beyond the
    // samples the population holds its last value)
    struct Schedule {
        std::vector<Real> v;
        Real t0 = 0, dt = 1;
        tg_bath_schedule desc(const Band<NDim>& b) const {
            tg_bath_schedule d{};
            d.samples = v.data();
            d.n_samples = (int32_t)v.size();
            d.batch = 1;
            d.t_start = t0;
            d.dt = dt;
            d.atol = b.absoluteErrorTolerance;
            d.rtol = b.relativeErrorTolerance;
            return d;
        }
    };
    static Real constrainedValue(const Band<NDim>& b, Real t) {
        return MaterialStatus<NDim>::uniformValue(b.constrainedPopulation(t), "constrained phonon population");
    }
    static Schedule sampleConstraint(const Band<NDim>& b, Real t0, Real t1, Real denseStep) {
        if (!b.constrainedPopulation) throw std::runtime_error("tortoise_gpu: constrained band without a constrainedPopulation");
        Schedule sc;
        sc.t0 = t0;
        sc.dt = denseStep / (Real)(b.constraintSamplesPerDenseStep > 0 ? b.constraintSamplesPerDenseStep : 1);
        const int64_t ns = (int64_t)std::ceil(2.0 * (t1 - t0) / sc.dt) + 2;
        if (ns > (int64_t)1 << 24) throw std::runtime_error("tortoise_gpu: too many constraint samples");
        sc.v.resize((size_t)ns);
        for (int64_t k = 0; k < ns; ++k) sc.v[(size_t)k] = constrainedValue(b, t0 + (Real)k * sc.dt);
        return sc;
    }
    // what the backend cannot honour is refused, not ignored
    void checkConstraints(const Band<NDim>& el, const Band<NDim>& pb) const {
        if (el.constrained)
            throw std::runtime_error("tortoise_gpu: a constrained electron band is not supported by the backend");
        if (pb.constrained && material->dynamicPhonons)
            throw std::runtime_error(
                "tortoise_gpu: a constrained phonon band cannot evolve (dynamicPhonons); use it as the bath");
        for (const auto& b : material->bands)
            if (b.constrained && b.name != el.name && b.name != pb.name)
                throw std::runtime_error("tortoise_gpu: constrained band " + b.name + " is not on the backend's channel");
    }

    // MaterialTimeStatus.cpp:102-137 with the reference operator: new MC points
    // (material->constructMCPoints()), then DP5(4) of the electron band's DG
    // coefficients on the device (tg_tuple_dp5).
    void propagateReference(Real finalTime, Real denseStep) {
        const MaterialStatus<NDim> start = valVec.back();
        tg_tuple_ctx* t = material->tupleBackend();
        material->constructMCPoints();
        const auto& ch = material->tupleChannel();
        const Band<NDim>& el = (*material)[ch.electronBand];
        const Band<NDim>& pb = (*material)[ch.phononBand];
        checkConstraints(el, pb);
        const bool dyn = material->dynamicPhonons;
        std::vector<Real> f = start[ch.electronBand].vec;
        const size_t ne = f.size();
        if (dyn) f.insert(f.end(), start[ch.phononBand].vec.begin(), start[ch.phononBand].vec.end());
        const int64_t n = (int64_t)f.size();
        const Real t0 = argVec.back();
        const Real bath = (dyn || pb.constrained) ? 0.0 : start.bathOccupation();
        int32_t flags = quiet ? TG_QUIET : 0;
        if (pb.constrained) {
            const Schedule sc = sampleConstraint(pb, t0, finalTime, denseStep);
            tg_bath_schedule d = sc.desc(pb);
            material->checkTuple(tg_tuple_set_bath_schedule(t, &d));
            flags |= TG_BATH_SCHEDULE;
        }
        const int maxDense = (int)std::floor((finalTime - t0) / denseStep + 1e-9) + 1;
        std::vector<Real> dense((size_t)maxDense * n);
        material->checkTuple(tg_tuple_set_amplitude(t, ch.amplitude));
        tg_status s = tg_tuple_dp5(t, f.data(), el.mesh->numberElements, 1, t0, finalTime, denseStep,
                                   el.relativeErrorTolerance, el.absoluteErrorTolerance,
                                   (dyn || pb.constrained) ? nullptr : &bath, flags, dense.data(), maxDense,
                                   &lastStats, nullptr, 0);
        if (s != TG_OK && s != TG_ERR_STEP_UNDERFLOW) material->checkTuple(s);
        if (s == TG_ERR_STEP_UNDERFLOW) std::printf("The step size has become too small to be handled\n");
        const int nd = lastStats.n_dense < maxDense ? lastStats.n_dense : maxDense;
        for (int k = 0; k < nd; ++k) {
            MaterialStatus<NDim> st = start;
            std::copy(dense.begin() + (size_t)k * n, dense.begin() + (size_t)k * n + ne,
                      st[ch.electronBand].vec.begin());
            if (dyn)
                std::copy(dense.begin() + (size_t)k * n + ne, dense.begin() + (size_t)(k + 1) * n,
                          st[ch.phononBand].vec.begin());
            const Real td = t0 + (k + 1) * denseStep;
            if (pb.constrained) st[ch.phononBand] = pb.constrainedPopulation(td);   // enforceValue at td
            argVec.push_back(td);
            valVec.push_back(std::move(st));
        }
    }
};

}  // namespace Tortoise
