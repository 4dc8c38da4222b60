// Golden-vector harness around the compiled reference solver -- TEST INFRASTRUCTURE ONLY.
//
// The reference (solver.cpp) is compiled from where it lies under the
// reference checkout (the path comes from -DREF_SOLVER_CPP); nothing of it is
// copied here.  Its `main` is renamed so that this file can drive the class:
//
//   * the constructor reads INP.DAT from the working directory, assembles the
//     operators and enters timeLoop(), which runs for MAXSTEP = 5e6 steps and
//     prints "loop dmax" once per step (solver.cpp:1889);
//   * std::cout's buffer is replaced by one that counts those lines; on the
//     line of step L the object's public arrays are written to
//     <out>/s<L>_<name>.bin (n1 rows of n0 doubles, xi fastest) and, on the
//     last requested step, the process exits.
//
// Optional runtime changes applied from the hook after step 0 (documented in
// the fixture provenance): the Richardson number (so that the buoyancy terms
// of steps >= 1 are exercised; Re, Pr and dt stay at the compiled values
// because the operators were assembled with them).  On the step named by
// --solve-step the hook also calls the reference's own sip9p()/gauss()
// repeatedly on copies of that step's systems to obtain their converged
// solutions.
//
// Usage: ref_harness <workdir-with-INP.DAT> <outdir> <last_step> [--ri X] [--solve-step S]
//        ref_harness <workdir-with-INP.DAT> <outdir> <steps> --time 1
//   (timing mode: nothing is dumped; the wall time of each of the first <steps>
//   steps after step 0 is printed, one per line, and the process exits)
#include <chrono>
#include <cstdio>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <streambuf>
#include <string>
#include <unistd.h>
#include <vector>

#define main cfd_reference_main
#include REF_SOLVER_CPP
#undef main

namespace {

Solver *g_s = nullptr;
std::string g_out;
int g_last = 2;
int g_solve_step = -1;
double g_ri = -1.0;
int g_line = 0;
bool g_time = false;
std::chrono::steady_clock::time_point g_t0;

void dump2(const char *tag, const char *name, double **a) {
    const int n0 = n[0], n1 = n[1];
    std::vector<double> buf((size_t)n0 * n1);
    for (int j = 0; j < n1; j++)
        for (int i = 0; i < n0; i++) buf[(size_t)j * n0 + i] = a[i][j];
    std::string p = g_out + "/" + tag + "_" + name + ".bin";
    FILE *f = fopen(p.c_str(), "wb");
    if (!f) { perror(p.c_str()); _exit(3); }
    fwrite(buf.data(), sizeof(double), buf.size(), f);
    fclose(f);
}

void dump1(const char *tag, const char *name, const double *a, int len) {
    std::string p = g_out + "/" + tag + "_" + name + ".bin";
    FILE *f = fopen(p.c_str(), "wb");
    if (!f) { perror(p.c_str()); _exit(3); }
    fwrite(a, sizeof(double), (size_t)len, f);
    fclose(f);
}

void dump_scalar(const char *tag, const char *name, double v) { dump1(tag, name, &v, 1); }

double **alloc_copy(double **src) {
    double **a = g_s->allocate2D(Solver::np1, Solver::np2);
    for (int i = 0; i < Solver::np1; i++)
        for (int j = 0; j < Solver::np2; j++) a[i][j] = src ? src[i][j] : 0.0;
    return a;
}

void dump_setup() {
    Solver &s = *g_s;
    const char *t = "setup";
    dump2(t, "x", s.x[0]); dump2(t, "y", s.x[1]);
    dump2(t, "xix", s.dxix); dump2(t, "xiy", s.dxiy); dump2(t, "etx", s.dex); dump2(t, "ety", s.dey);
    dump2(t, "alph", s.alph); dump2(t, "beta", s.beta); dump2(t, "gamma", s.gamma);
    dump2(t, "ajac", s.ajac);
    // 9-point p' operator
    dump2(t, "ap", s.ap); dump2(t, "ae", s.ae); dump2(t, "aw", s.aw); dump2(t, "an", s.an);
    dump2(t, "as", s.as); dump2(t, "ane", s.ane); dump2(t, "anw", s.anw); dump2(t, "ase", s.ase);
    dump2(t, "asw", s.asw);
    // momentum operator (25 planes)
    double **au[] = {s.aup, s.aue, s.auw, s.aun, s.aus, s.aune, s.aunw, s.ause, s.ausw,
                     s.auee, s.auww, s.aunn, s.auss, s.aunnee, s.aunnww, s.aussee, s.aussww,
                     s.aunne, s.aunnw, s.ausse, s.aussw, s.aunee, s.aunww, s.ausee, s.ausww};
    double **at[] = {s.atp, s.ate, s.atw, s.atn, s.ats, s.atne, s.atnw, s.atse, s.atsw,
                     s.atee, s.atww, s.atnn, s.atss, s.atnnee, s.atnnww, s.atssee, s.atssww,
                     s.atnne, s.atnnw, s.atsse, s.atssw, s.atnee, s.atnww, s.atsee, s.atsww};
    const char *nm[] = {"p", "e", "w", "n", "s", "ne", "nw", "se", "sw",
                        "ee", "ww", "nn", "ss", "nnee", "nnww", "ssee", "ssww",
                        "nne", "nnw", "sse", "ssw", "nee", "nww", "see", "sww"};
    for (int k = 0; k < 25; k++) {
        std::string a = std::string("au") + nm[k], b = std::string("at") + nm[k];
        dump2(t, a.c_str(), au[k]);
        dump2(t, b.c_str(), at[k]);
    }
    const int n0 = n[0];
    dump1(t, "bus", s.bus, n0); dump1(t, "buse", s.buse, n0); dump1(t, "busw", s.busw, n0);
    dump1(t, "bun", s.bun, n0); dump1(t, "bune", s.bune, n0); dump1(t, "bunw", s.bunw, n0);
    dump1(t, "bts", s.bts, n0); dump1(t, "btse", s.btse, n0); dump1(t, "btsw", s.btsw, n0);
    dump1(t, "btn", s.btn, n0); dump1(t, "btne", s.btne, n0); dump1(t, "btnw", s.btnw, n0);
    dump1(t, "xnox", s.xnox, n0); dump1(t, "xnoy", s.xnoy, n0);
    double sc[8] = {s.Re, s.Pr, s.dt, s.Ri, s.uinf, s.vinf, s.ar, s.dxi[0]};
    dump1(t, "scalars", sc, 8);
}

void dump_state(int L) {
    Solver &s = *g_s;
    char t[32];
    snprintf(t, sizeof t, "s%d", L);
    dump2(t, "u", s.u[0]); dump2(t, "v", s.u[1]); dump2(t, "T", s.u[2]);
    dump2(t, "up", s.up[0]); dump2(t, "vp", s.up[1]);
    dump2(t, "us", s.us[0]); dump2(t, "vs", s.us[1]);
    dump2(t, "uold", s.uold[0]); dump2(t, "vold", s.uold[1]); dump2(t, "Told", s.uold[2]);
    dump2(t, "p", s.p); dump2(t, "pcor", s.pcor); dump2(t, "q", s.q);
    dump2(t, "qu", s.qu); dump2(t, "qv", s.qv); dump2(t, "qt", s.qt);
    dump2(t, "qup", s.qup); dump2(t, "qvp", s.qvp);
    dump2(t, "dil", s.dil); dump2(t, "uxi", s.uxi); dump2(t, "uet", s.uet);
    dump2(t, "sol", s.sol);
    dump2(t, "si", s.si); dump2(t, "vort", s.vort);
    dump_scalar(t, "dmax", s.dmax);
    dump_scalar(t, "Ri", s.Ri);
}

// Converged solutions of step L's p' and u systems with the reference's own
// solvers, called repeatedly: each call cuts the L1 residual of its own start
// by 0.75e-2, so 7 calls reach ~1e-15 relative.  (A call that starts at the
// rounding floor runs to the 1e5-iteration cap, which is why the count is
// fixed and small.)
void dump_converged(int L) {
    Solver &s = *g_s;
    char t[32];
    snprintf(t, sizeof t, "s%d", L);
    double **phi = alloc_copy(nullptr);
    for (int k = 0; k < 7; k++)
        s.sip9p(s.ap, s.ae, s.as, s.an, s.aw, s.ase, s.asw, s.ane, s.anw, phi, s.q);
    dump2(t, "pcor_conv", phi);
    // one reference call from zero: equals the interior of pcor before its BCs
    double **one = alloc_copy(nullptr);
    s.sip9p(s.ap, s.ae, s.as, s.an, s.aw, s.ase, s.asw, s.ane, s.anw, one, s.q);
    dump2(t, "pcor_one", one);
    // u system of this step: A_u us = qu, started from the step's initial u (uold)
    double **sol = alloc_copy(s.uold[0]);
    for (int k = 0; k < 6; k++)
        s.gauss(s.aup, s.aue, s.aus, s.aun, s.auw, s.ause, s.ausw, s.aune, s.aunw, s.auss,
                s.aussee, s.aussww, s.ausse, s.aussw, s.ausee, s.ausww, s.aunn, s.aunnee,
                s.aunnww, s.aunne, s.aunnw, s.aunee, s.aunww, s.auee, s.auww, sol, s.qu);
    dump2(t, "u_conv", sol);
    // T system as the reference calls it (the (-1,+1) plane in the (-1,+2) slot)
    double **solt = alloc_copy(s.uold[2]);
    for (int k = 0; k < 6; k++)
        s.gauss(s.atp, s.ate, s.ats, s.atn, s.atw, s.atse, s.atsw, s.atne, s.atnw, s.atss,
                s.atssee, s.atssww, s.atsse, s.atssw, s.atsee, s.atsww, s.atnn, s.atnnee,
                s.atnnww, s.atnne, s.atnw, s.atnee, s.atnww, s.atee, s.atww, solt, s.qt);
    dump2(t, "T_conv", solt);
    s.deallocate2D(phi, Solver::np1);
    s.deallocate2D(one, Solver::np1);
    s.deallocate2D(sol, Solver::np1);
    s.deallocate2D(solt, Solver::np1);
}

void copy_text(const char *src, const char *dst_name) {
    FILE *f = fopen(src, "rb");
    if (!f) return;
    std::string p = g_out + "/" + dst_name;
    FILE *o = fopen(p.c_str(), "wb");
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) fwrite(buf, 1, k, o);
    fclose(f);
    fclose(o);
}

void on_step_line(int L) {
    if (g_time) {
        const auto t = std::chrono::steady_clock::now();
        if (L >= 1) fprintf(stderr, "%.3f\n", std::chrono::duration<double, std::milli>(t - g_t0).count());
        if (L >= g_last) {
            fflush(nullptr);
            _exit(0);
        }
        g_t0 = std::chrono::steady_clock::now();
        return;
    }
    if (L == 0) {
        dump_setup();
        // the momentum-solver inputs of step 0 start from the initial fields; keep them
    }
    dump_state(L);
    if (L == g_solve_step) dump_converged(L);
    if (L == 0 && g_ri >= 0.0) g_s->Ri = g_ri;
    if (L == 1) {
        // step 0 wrote its coefficient history (Nu) and local dT/dn profile
        copy_text("COEFF_HIS.dat", "s0_COEFF_HIS.dat");
        copy_text("COEFF_HIS_pr_vor.dat", "s0_COEFF_HIS_pr_vor.dat");
        copy_text("SURF_DIST.dat", "s0_SURF_DIST.dat");
        copy_text("spt100.dat", "s0_spt100.dat");
    }
    if (L >= g_last) {
        fflush(nullptr);
        _exit(0);
    }
}

class LineHook : public std::streambuf {
  protected:
    int overflow(int c) override {
        if (c == '\n') {
            int L = g_line++;
            on_step_line(L);
        }
        return c == EOF ? 0 : c;
    }
};

}  // namespace

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s workdir outdir last_step [--ri X] [--solve-step S]\n", argv[0]);
        return 2;
    }
    {
        char *abs = realpath(argv[2], nullptr);
        if (!abs) { perror(argv[2]); return 2; }
        g_out = abs;
        free(abs);
    }
    g_last = atoi(argv[3]);
    for (int a = 4; a + 1 < argc; a += 2) {
        if (!strcmp(argv[a], "--ri")) g_ri = atof(argv[a + 1]);
        else if (!strcmp(argv[a], "--solve-step")) g_solve_step = atoi(argv[a + 1]);
        else if (!strcmp(argv[a], "--time")) g_time = atoi(argv[a + 1]) != 0;
    }
    if (chdir(argv[1]) != 0) { perror("chdir"); return 2; }
    static LineHook hook;
    std::cout.rdbuf(&hook);
    void *mem = ::operator new(sizeof(Solver));
    g_s = static_cast<Solver *>(mem);
    new (mem) Solver();  // never returns: timeLoop() runs until the hook exits
    return 1;
}
