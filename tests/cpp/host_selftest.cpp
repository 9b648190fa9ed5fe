// CPU checks of the Tortoise:: host layer (no GPU calls): mesh numbering,
// DG <-> nodal conversion, Function arithmetic and integrals against closed
// forms.  Prints "ok" and exits 0 on success.
#include <TORTOISE>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

using namespace Tortoise;

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);     \
            ++fails;                                                    \
        }                                                               \
    } while (0)

int main() {
    const Real pi = M_PI;
    Region<2> zone(Point<2>({-pi, -pi}), {Point<2>({2 * pi, 0.0}), Point<2>({0.0, 2 * pi})});
    Mesh<2> m(zone, Point<2>({0., 0.}), Point<2>({1., 1.}), {6, 5});
    CHECK(m.numberElements == 60 && m.numberNodes() == 42);
    // Mesh.cpp:55-87 vertex order of section (0,0)
    CHECK(m.elementNode(0, 0) == 0 && m.elementNode(0, 1) == 1 && m.elementNode(0, 2) == 8);
    CHECK(m.elementNode(1, 0) == 0 && m.elementNode(1, 1) == 7 && m.elementNode(1, 2) == 8);
    // a linear field round-trips exactly and integrates in closed form
    Function<2> lin(m, [](Point<2> k) { return 1.0 + 0.5 * k(0) - 0.25 * k(1); });
    auto nodal = lin.toNodal();
    Function<2> back = Function<2>::fromNodal(m, nodal);
    Real d = 0;
    for (size_t i = 0; i < lin.vec.size(); ++i) d = std::fmax(d, std::fabs(lin.vec[i] - back.vec[i]));
    CHECK(d < 1e-14);
    CHECK(std::fabs(lin.integrate() - 4 * pi * pi) < 1e-12);   // odd parts vanish on the box
    // int x^2 over [-pi,pi]^2 = (2pi)(2pi^3/3); exact for P1*P1
    Function<2> x(m, [](Point<2> k) { return k(0); });
    CHECK(std::fabs(x.integrate(x) - 2 * pi * 2 * pi * pi * pi / 3) < 1e-10);
    // nodal product and constant shift
    Function<2> one(m, 1.0);
    Function<2> sq = x * x;
    CHECK(std::fabs(sq.integrate() - (x.integrate(x) + 0.0)) > 0.0);  // lumped != exact
    Function<2> y = one;
    y += 2.0;
    CHECK(std::fabs(y.integrate() - 3 * 4 * pi * pi) < 1e-11);
    // Band / Material bookkeeping
    Material<2> mat("m", zone);
    mat.addBand("el:C", -1, fermion, m, [](Point<2> k) { return 0.5 * k.squaredNorm(); });
    mat.addBand("pn:o", 0, boson, m, 0.05);
    auto ch = mat.addScatteringChannelsWildcards<3>({"el:*", "pn:*", "el:*"}, {in, in, out}, 0);
    CHECK(ch.size() == 1 && ch[0]->electronBand == "el:C" && ch[0]->phononBand == "pn:o");
    // the 4-leg e-e channel on the same fermion band (the phonon band never matches)
    auto ee = mat.addScatteringChannelsWildcards<4>({"el:*", "el:*", "el:*", "*"}, {in, in, out, out}, 0);
    CHECK(ee.size() == 1 && ee[0]->legs == 4 && ee[0]->electronBand == "el:C");
    CHECK(mat.channels.size() == 2);
    MaterialStatus<2> s(mat);
    s.setToEquilibrium(0.02, 0.5, 0.0);
    CHECK(std::fabs(s.bathOccupation() - 1.0 / std::expm1(0.05 / 0.02)) < 1e-15);
    // a non-uniform phonon population is refused, not replaced by one value
    MaterialStatus<2> s2 = s;
    s2["pn:o"] = Function<2>(m, [](Point<2> k) { return 0.1 + 0.01 * k(0); });
    bool threw = false;
    try { (void)s2.bathOccupation(); } catch (const std::runtime_error&) { threw = true; }
    CHECK(threw);
    // a constrained electron band is refused before any device work
    mat["el:C"].constrained = true;
    mat["el:C"].constrainedPopulation = [&m](Real) { return Function<2>(m, 0.5); };
    MaterialTimeStatus<2> ts(s);
    threw = false;
    try { ts.propagate(1.0, 0.1); } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find("constrained electron band") != std::string::npos;
    }
    CHECK(threw);
    mat["el:C"].constrained = false;
    // the hexagonal zone of the reference's example (examples/main.cpp:37-39):
    // a skewed region, a sub-mesh of it, node positions, element volume, and
    // the W-form backend's refusal (it meshes axis-aligned zones)
    const Real h3 = std::sqrt(3.0) / 2.0;
    Region<2> hex(Point<2>({0.0, 0.0}), {Point<2>({1.0, 0.0}), Point<2>({-0.5, h3})});
    CHECK(!hex.axisAligned());
    Mesh<2> hm(hex, Point<2>({0.1, 0.1}), Point<2>({0.3, 0.3}), {3, 3});
    CHECK(!hm.axisAligned);
    // node (i, j) = origin + (i/3) side0 + (j/3) side1, origin = 0.1 (g0 + g1), side_a = 0.3 g_a
    const Point<2> o = hm.node(0), p = hm.node(2 + 4 * 1);
    CHECK(std::fabs(o(0) - 0.1 * (1.0 - 0.5)) < 1e-15 && std::fabs(o(1) - 0.1 * h3) < 1e-15);
    CHECK(std::fabs(p(0) - (o(0) + 0.2 * 1.0 + 0.1 * -0.5)) < 1e-15 && std::fabs(p(1) - (o(1) + 0.1 * h3)) < 1e-15);
    CHECK(std::fabs(hm.elemVolume - 0.09 * h3 / 9.0 / 2.0) < 1e-16);   // |det(0.3 g)| / 9 / 2
    // the integral of 1 over the sub-mesh is its area
    CHECK(std::fabs(Function<2>(hm, 1.0).integrate() - 0.09 * h3) < 1e-14);
    Material<2> hmat("hex", hex);
    hmat.addBand("el:C", -1, fermion, hm, [](Point<2> k) { return 0.5 * k.squaredNorm(); });
    hmat.addBand("pn:o", 0, boson, hm, 0.05);
    hmat.addScatteringChannelsWildcards<3>({"el:*", "pn:*", "el:*"}, {in, in, out}, 1);
    threw = false;
    try { (void)hmat.backend(0.1, 1); } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find("axis-aligned") != std::string::npos;
    }
    CHECK(threw);
    if (fails == 0) std::printf("ok\n");
    return fails ? 1 : 0;
}
