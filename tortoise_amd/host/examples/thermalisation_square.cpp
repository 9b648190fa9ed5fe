// Hot-electron thermalisation on a square zone, written against the
// Tortoise:: host API the way the reference's tutorials are (compare
// tutorials/TORTOISE_Workshop/Thermalisation/main.cpp): one parabolic
// electron band, a dispersionless optical phonon band held as a bath, the
// e-ph channel, optionally the e-e channel on the electron band, DP5 time
// propagation with dense output, and the number / energy sums of every dense
// step.  Runs the collision operator on the GPU.
//
// With ref=1 the collision operator is the reference's own: element tuples
// and the Monte-Carlo delta-shell integral per tuple on the DG coefficients.
//
// With ref=1 dyn=1 the phonon band evolves instead of acting as a bath.
//
// With pulse=A > 0 the phonon band is constrained to a pulsed bath,
// n(t) = n_BE + A (1 - x^2)^2 for |x| < 1, x = (t - t_final/4) / (t_final/10),
// the way the reference's examples/main.cpp:93-94 drives a band
// (Band::constrained and constrainedPopulation).
//
// usage: thermalisation_square [sections=32] [t_final=100] [quiet=1] [g_ee=0] [ref=0] [dyn=0] [pulse=0]
#include <TORTOISE>

#include <cmath>
#include <cstdio>
#include <cstdlib>

using namespace Tortoise;

int main(int argc, char** argv) {
    const int res = argc > 1 ? std::atoi(argv[1]) : 32;
    const Real tFinal = argc > 2 ? std::atof(argv[2]) : 100.0;
    const bool quiet = argc > 3 ? std::atoi(argv[3]) != 0 : true;
    const Real gEE = argc > 4 ? std::atof(argv[4]) : 0.0;
    const bool ref = argc > 5 ? std::atoi(argv[5]) != 0 : false;
    const bool dynPh = argc > 6 ? std::atoi(argv[6]) != 0 : false;
    const Real pulse = argc > 7 ? std::atof(argv[7]) : 0.0;
    const Real pi = M_PI;

    Region<2> zone(Point<2>({-pi, -pi}), {Point<2>({2 * pi, 0.0}), Point<2>({0.0, 2 * pi})});
    Mesh<2> mesh(zone, Point<2>({0., 0.}), Point<2>({1., 1.}), {res, res});
    Material<2> mat("SquareParabolic", zone);
    mat.addBand("el:C", -1, fermion, mesh, [](Point<2> k) { return 0.5 * k.squaredNorm(); });
    mat.addBand("pn:opt", 0, boson, mesh, 0.05);
    mat.referenceOperator = ref;
    mat.dynamicPhonons = dynPh;
    for (auto sc : mat.addScatteringChannelsWildcards<3>({"el:*", "pn:*", "el:*"}, {in, in, out}, 1)) {
        sc->amplitude = 0.1 * 0.1;      // g^2
        sc->deltaBroadening = 0.01;
    }
    if (gEE != 0.0)
        for (auto sc : mat.addScatteringChannelsWildcards<4>({"el:*", "el:*", "el:*", "el:*"},
                                                             {in, in, out, out}, 0)) {
            sc->amplitude = gEE * gEE;  // g_ee^2 in front of V(q)^2
            sc->screening = 0.09;       // kappa^2
            sc->deltaBroadening = 0.05;
        }
    mat["el:C"].relativeErrorTolerance = 1e-8;
    mat["el:C"].absoluteErrorTolerance = 1e-12;

    const Real T = 0.02, mu = 0.5;
    MaterialStatus<2> start(mat);
    start.setToEquilibrium(T, mu, 0.0);   // phonon bath at n_BE(omega0, T)
    start["el:C"] += Function<2>(mesh, [](Point<2> k) {
        const Real e = 0.5 * k.squaredNorm();
        return 0.2 * std::exp(-0.5 * std::pow((e - 0.9) / 0.03, 2));
    });

    if (ref) {
        // relaxation rates of the initial state (Material::scatteringRates, as
        // the reference's examples/main.cpp:84-87 prints them)
        Function<2> rates = mat.scatteringRates<3>({"el:*", "pn:*", "el:*"}, {in, in, out}, start, "el:C");
        std::printf("rates integral %.15e\n", rates.integrate());
    }
    if (pulse > 0.0) {
        const Real nBE = 1.0 / std::expm1(0.05 / T), tp = tFinal / 4, w = tFinal / 10;
        mat["pn:opt"].constrained = true;
        mat["pn:opt"].constrainedPopulation = [&mesh, nBE, tp, w, pulse](Real t) {
            const Real x = (t - tp) / w;
            const Real q = 1.0 - x * x;
            return Function<2>(mesh, nBE + (std::fabs(x) < 1.0 ? pulse * q * q : 0.0));
        };
    }
    MaterialTimeStatus<2> dyn(start);
    dyn.quiet = quiet;
    dyn.propagateDeterministic(tFinal, tFinal / 100.0);

    std::printf("steps %lld accepted %lld rejected %lld\n", (long long)dyn.lastStats.attempted,
                (long long)dyn.lastStats.accepted, (long long)dyn.lastStats.rejected);
    std::printf("t | electron number | electron energy\n");
    for (int i = 0; i < dyn.size(); i += dyn.size() > 10 ? dyn.size() / 10 : 1)
        std::printf("%.4f | %.15e | %.15e\n", dyn.arg(i), dyn[i]["el:C"].integrate(),
                    dyn[i]["el:C"].integrate(mat["el:C"].dispersion));
    const Function<2>& fEnd = dyn.back()["el:C"];
    std::printf("final %.4f | %.15e | %.15e\n", dyn.argback(), fEnd.integrate(),
                fEnd.integrate(mat["el:C"].dispersion));
    if (pulse > 0.0) std::printf("bath at the end %.15e\n", dyn.back().bathOccupation());
    if (mat.dynamicPhonons) {
        const Function<2>& pEnd = dyn.back()["pn:opt"];
        std::printf("phonons %.15e | %.15e\n", pEnd.integrate(), pEnd.integrate(mat["pn:opt"].dispersion));
    }
    return 0;
}
