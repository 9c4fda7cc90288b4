// tcp_run -- standalone driver of the MD loop of main.f90:81-170 on the GPU
// path: reads tcp.prm and restart.xyz from the current directory and writes
// ener##.dat, RDF###.dat, VDF###.dat and REST##.xyz with the bytes the
// reference's result.f90 writes for the same run (the Fortran drop-in run as
// `mpirun -np 2` produces identical files in FP64).
//
//   tcp_run                  (TCP_PRECISION=MIXED selects the mixed pair kernel)
//
// Between output events the steps run on the device in one tcp_md_steps call;
// the Langevin thermostat (STOCH) takes one call per step with the noise seed
// sequence of the Fortran drop-in.  Only restart inputs (rest = 'ON') are
// supported: the reference's fresh start draws velocities from the compiler's
// unseeded RANDOM_NUMBER (init.f90:45-68).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tcp_gpu.h"
#include "tcp_io.h"

namespace {

constexpr double kTps = 10.180505306645389;  // data.f90:22
constexpr double kEps = 14.399644154020082;  // data.f90:21
constexpr int kNdist = 500;                  // data.f90:16
constexpr int kNref = 0;                     // data.f90:16
constexpr uint64_t kStochSeed = 20240607;    // the drop-in's first seed minus one

int die(const std::string& m)
{
    std::fprintf(stderr, "tcp_run: %s\n", m.c_str());
    return 1;
}

int thermo_code(const char* s)
{
    if (!std::strncmp(s, "BEREND", 6)) return TCP_BEREND;
    if (!std::strncmp(s, "ISOKIN", 6)) return TCP_ISOKIN;
    if (!std::strncmp(s, "STOCH", 5)) return TCP_STOCH;
    return TCP_NOTEMP;
}

// Fortran NINT: nearest, halves away from zero
long nint(double v) { return std::lround(v); }

}  // namespace

int main()
{
    tcp_prm prm;
    char err[256] = {0};
    if (tcp_io_read_prm("tcp.prm", &prm, err)) return die(err);
    if (!std::strncmp(prm.rest, "OFF", 3))
        return die("rest = 'OFF' needs the reference's unseeded velocity draw; start from restart.xyz");
    // main.f90:230-236 (times to internal units), 55-61 (step counts)
    const double tmax = prm.tmax / kTps, trest = prm.trest / kTps, tdump = prm.tdump / kTps,
                 tsamp = prm.tsamp / kTps, dt = prm.dt / kTps;
    const long ndump = nint(tdump / dt), nrest = nint(trest / dt), nsamp = nint(tsamp / dt);
    const long nloop_max = nint(tmax / dt);
    const int thermo = thermo_code(prm.thermo);
    const int n = prm.NTel + prm.NTion;

    tcp_params p{};
    p.ne = prm.NTel;
    p.ni = prm.NTion;
    for (int d = 0; d < 3; ++d) {
        p.K[d] = prm.K[d];
        p.box[d] = prm.box[d];
    }
    p.order = 4;  // Force_SPME4
    p.recip = TCP_RECIP_SPME;
    p.flags = 0;
    p.q[0] = prm.q[0];
    p.q[1] = prm.q[1];
    p.xm[0] = prm.xm[0];
    p.xm[1] = prm.xm[1];
    p.rc = prm.rc;
    p.rs = prm.rs;
    p.alpha = 0.0;
    p.eps = kEps;
    p.dx_ee = prm.dx[0];
    p.dx_ei = prm.dx[1];
    p.dx_ii = prm.dx[2];
    p.dv_e = prm.Te * 4.0 / (double)kNdist;  // main.f90:286-287
    p.dv_i = prm.Ti * 4.0 / (double)kNdist;
    const char* prec = std::getenv("TCP_PRECISION");
    const int precision = prec && (!std::strncmp(prec, "MIXED", 5) || !std::strncmp(prec, "mixed", 5))
                              ? TCP_MIXED : TCP_FP64;

    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n);
    std::vector<int> id(n);
    if (tcp_io_read_restart("restart.xyz", n, x.data(), v.data(), f.data(), id.data(), err))
        return die(err);
    tcp_ctx* ctx = nullptr;
    if (int rc = tcp_create(&ctx, &p, precision)) {
        if (rc == TCP_ERR_CUTOFF) std::printf(" *** error at cut-off radius *** \n");
        return die(tcp_last_error(nullptr));
    }
    auto check = [&](int rc, const char* what) {
        if (rc) {
            std::fprintf(stderr, "tcp_run: %s failed: %s\n", what, tcp_last_error(ctx));
            std::exit(1);
        }
    };
    check(tcp_set_state(ctx, x.data(), v.data(), f.data()), "tcp_set_state");

    std::string ener = "ener00.dat";
    if (tcp_io_ener_open(ener.c_str())) return die("cannot write " + ener);
    const double rel = (double)(prm.NTel * 3 - kNref), rion = (double)(prm.NTion * 3 - kNref);
    std::vector<double> trace, rdf(3 * (kNdist + 1)), vdf(2 * (kNdist + 1));
    double tnow = 0.0;
    long nloop = 0;
    auto due = [](long k, long every) { return every > 0 && k % every == 0; };
    const auto t0 = std::chrono::steady_clock::now();
    while (nloop < nloop_max) {
        // steps up to the next output event (or the end)
        long k = 1;
        if (thermo != TCP_STOCH)
            while (nloop + k < nloop_max && !due(nloop + k, ndump) && !due(nloop + k, nsamp) &&
                   !due(nloop + k, nrest))
                ++k;
        trace.assign(6 * (size_t)k, 0.0);
        const uint64_t seed = kStochSeed + (uint64_t)nloop + 1;
        check(tcp_md_steps(ctx, (int)k, thermo, dt, prm.Te, prm.Ti, prm.tau, prm.alpha, seed,
                           trace.data()), "tcp_md_steps");
        for (long s = 0; s < k; ++s) tnow = tnow + dt;  // main.f90:83
        nloop += k;
        const double* row = trace.data() + 6 * (size_t)(k - 1);  // Ur Uo Um Epot mv2el mv2ion
        if (due(nloop, ndump)) {  // DUMP, result.f90:210-268
            std::fill(rdf.begin(), rdf.end(), 0.0);
            std::fill(vdf.begin(), vdf.end(), 0.0);
            check(tcp_get_histograms(ctx, rdf.data(), vdf.data(), 1), "tcp_get_histograms");
            if (tcp_io_write_dump((int)(nloop / ndump), prm.NTel, prm.NTion, (int)ndump, prm.box, prm.dx,
                                  p.dv_e, p.dv_i, rdf.data(), vdf.data()))
                return die("cannot write the RDF/VDF files");
        }
        if (due(nloop, nsamp)) {  // SAMPLING, result.f90:290-326
            const double epot = row[2] + row[3];
            const double etot = 0.5 * (row[4] + row[5]) + epot;
            if (tcp_io_ener_row(ener.c_str(), tnow * kTps, row[4] / rel, row[5] / rion, epot, etot))
                return die("cannot write " + ener);
        }
        if (due(nloop, nrest)) {  // Restart, result.f90:3-134
            const int nfreq = (int)(nloop / nrest);
            check(tcp_get_state(ctx, x.data(), v.data(), f.data()), "tcp_get_state");
            if (tcp_io_write_rest(nfreq, prm.NTel, prm.NTion, x.data(), v.data(), f.data(), id.data()))
                return die("cannot write the restart file");
            char name[32];
            std::snprintf(name, sizeof name, "ener%02d.dat", nfreq);
            ener = name;
            if (tcp_io_ener_open(ener.c_str())) return die("cannot write " + ener);
        }
    }
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("tcp_run: %ld steps of %d particles in %.3f s (%.3f ms/step)\n", nloop, n, wall,
                nloop ? wall / (double)nloop * 1e3 : 0.0);
    tcp_destroy(ctx);
    return 0;
}
