// Op-counting build of the oracle (measurement infrastructure, not product code).
//
// The oracle's C sources are compiled here as C++ with `double` replaced by a scalar
// type that counts every floating-point operation of the restated reference algorithm:
//   add, sub, mul, div, sqrt = 1; FMA = 2 (none occur: the reference is built without
//   contraction); exp, log, sin, cos, tan, atan, atan2, asin = 20; pow = 40;
//   negation, comparisons, fabs, fmin, fmax, round, fmod = 0.
// Expressions of literals only are folded by the compiler and not counted, so the
// count is the per-evaluation work.
//
// Usage: opcount <sites.txt> <nk1> <step_limit> <internal> <external> [parmod...]
//   sites.txt: one "r0 lat lon" line per site (geocentric, vertical incidence)
//   COR_OPCOUNT_SH (environment, optional): a file holding an SH model table -- int32
//   nm, then nm*nm doubles g and nm*nm doubles h (row n, column m) -- traced as the
//   crustal model of the set (lib/chaos/chaos.c:118-227)
// Output: one line per site: "site n sum_steps sum_attempts sum_ops field_evals"
// plus a total line; trajectories are rigidities 0.01..20 GV in 0.01 GV steps.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <type_traits>
#include <vector>

static thread_local unsigned long long g_ops = 0;

struct cdouble {
    double v;
    cdouble() = default;
    constexpr cdouble(double x) : v(x) {}
    explicit operator int() const { return (int)v; }
    explicit operator unsigned() const { return (unsigned)v; }
    explicit operator long() const { return (long)v; }
    explicit operator unsigned long() const { return (unsigned long)v; }
    explicit operator float() const { return (float)v; }
    cdouble operator-() const { return cdouble(-v); }
    cdouble operator+() const { return *this; }
    cdouble &operator+=(cdouble o) { ++g_ops; v += o.v; return *this; }
    cdouble &operator-=(cdouble o) { ++g_ops; v -= o.v; return *this; }
    cdouble &operator*=(cdouble o) { ++g_ops; v *= o.v; return *this; }
    cdouble &operator/=(cdouble o) { ++g_ops; v /= o.v; return *this; }
};

template <class T>
using arith = std::enable_if_t<std::is_arithmetic<T>::value, int>;

#define COR_BINOP(OP)                                                                     \
    inline cdouble operator OP(cdouble a, cdouble b) { ++g_ops; return cdouble(a.v OP b.v); } \
    template <class T, arith<T> = 0>                                                      \
    inline cdouble operator OP(cdouble a, T b) { ++g_ops; return cdouble(a.v OP (double)b); } \
    template <class T, arith<T> = 0>                                                      \
    inline cdouble operator OP(T a, cdouble b) { ++g_ops; return cdouble((double)a OP b.v); }
COR_BINOP(+)
COR_BINOP(-)
COR_BINOP(*)
COR_BINOP(/)
#undef COR_BINOP

#define COR_CMP(OP)                                                               \
    inline bool operator OP(cdouble a, cdouble b) { return a.v OP b.v; }          \
    template <class T, arith<T> = 0>                                              \
    inline bool operator OP(cdouble a, T b) { return a.v OP (double)b; }          \
    template <class T, arith<T> = 0>                                              \
    inline bool operator OP(T a, cdouble b) { return (double)a OP b.v; }
COR_CMP(<)
COR_CMP(>)
COR_CMP(<=)
COR_CMP(>=)
COR_CMP(==)
COR_CMP(!=)
#undef COR_CMP

inline cdouble sqrt(cdouble a) { ++g_ops; return cdouble(::sqrt(a.v)); }
#define COR_T20(F) \
    inline cdouble F(cdouble a) { g_ops += 20; return cdouble(::F(a.v)); }
COR_T20(exp)
COR_T20(log)
COR_T20(sin)
COR_T20(cos)
COR_T20(tan)
COR_T20(atan)
COR_T20(asin)
#undef COR_T20
inline cdouble atan2(cdouble a, cdouble b) { g_ops += 20; return cdouble(::atan2(a.v, b.v)); }
template <class T, arith<T> = 0>
inline cdouble pow(cdouble a, T b) { g_ops += 40; return cdouble(::pow(a.v, (double)b)); }
inline cdouble pow(cdouble a, cdouble b) { g_ops += 40; return cdouble(::pow(a.v, b.v)); }
inline cdouble fabs(cdouble a) { return cdouble(::fabs(a.v)); }
inline cdouble fmin(cdouble a, cdouble b) { return cdouble(::fmin(a.v, b.v)); }
inline cdouble fmax(cdouble a, cdouble b) { return cdouble(::fmax(a.v, b.v)); }
template <class T, arith<T> = 0>
inline cdouble fmax(cdouble a, T b) { return cdouble(::fmax(a.v, (double)b)); }
template <class T, arith<T> = 0>
inline cdouble fmax(T a, cdouble b) { return cdouble(::fmax((double)a, b.v)); }
inline cdouble round(cdouble a) { return cdouble(::round(a.v)); }
inline cdouble fmod(cdouble a, cdouble b) { return cdouble(::fmod(a.v, b.v)); }
template <class T, arith<T> = 0>
inline cdouble fmod(cdouble a, T b) { return cdouble(::fmod(a.v, (double)b)); }

#define _Thread_local thread_local
#define double cdouble
#include "../o_core.c"
#include "../o_t96.c"
#include "../o_ts05.c"
#include "../o_chaos.c"
#include "../o_igrf14_table.c"
#undef double

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: opcount sites.txt nk1 step_limit internal external [parmod...]\n");
        return 2;
    }
    std::vector<double> r0s, lats, lons;
    FILE *f = fopen(argv[1], "r");
    if (!f) return 2;
    double a, b, c;
    while (fscanf(f, "%lf %lf %lf", &a, &b, &c) == 3) {
        r0s.push_back(a);
        lats.push_back(b);
        lons.push_back(c);
    }
    fclose(f);
    const int nk1 = atoi(argv[2]);
    const uint64_t step_limit = strtoull(argv[3], nullptr, 10);
    o_models m;
    memset(&m, 0, sizeof m);
    m.internal = atoi(argv[4]);
    m.external = atoi(argv[5]);
    std::vector<cdouble> sh_g, sh_h;
    if (const char *sh = getenv("COR_OPCOUNT_SH")) {
        FILE *fs = fopen(sh, "rb");
        int32_t nm = 0;
        if (!fs || fread(&nm, sizeof nm, 1, fs) != 1 || nm < 2 || nm > 200) return 2;
        std::vector<double> g((size_t)nm * nm), h((size_t)nm * nm);
        if (fread(g.data(), sizeof(double), g.size(), fs) != g.size() ||
            fread(h.data(), sizeof(double), h.size(), fs) != h.size())
            return 2;
        fclose(fs);
        sh_g.assign(g.begin(), g.end());
        sh_h.assign(h.begin(), h.end());
        m.crustal = O_CRU_CHAOS;
        m.chaos_nm_max = nm;
        m.chaos_g = sh_g.data();
        m.chaos_h = sh_h.data();
    }
    o_params p;
    memset((void *)&p, 0, sizeof p);
    o_igrf14_recalc(2020, 1, 1, 1, 0, 0, 0, &p);
    for (int k = 6; k < argc && k - 6 < 10; k++) p.PARMOD[k - 6] = atof(argv[k]);
    unsigned long long tot_steps = 0, tot_att = 0, tot_ops = 0, tot_traj = 0;
    for (size_t s = 0; s < r0s.size(); s++) {
        o_scan sc;
        memset((void *)&sc, 0, sizeof sc);
        sc.r0 = r0s[s];
        sc.the0 = lats[s];
        sc.fi0 = lons[s];
        sc.the1 = lats[s];
        sc.fi1 = lons[s];
        sc.rig = 0.01;
        sc.del = 0.01;
        sc.rk = 20.0;
        sc.zn = -1.0;
        sc.nk1 = nk1;
        const int n = o_scan_count(&sc);
        unsigned long long s_steps = 0, s_att = 0, s_ops = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : s_steps, s_att, s_ops)
        for (int i = 0; i < n; i++) {
            o_traj t;
            g_ops = 0;
            o_trace_one(&sc, i, &m, &p, step_limit, 0.01, O_M_P, &t);
            s_steps += t.steps;
            s_att += t.attempts;
            s_ops += g_ops;
        }
        printf("site %zu n %d steps %llu attempts %llu ops %llu\n", s, n, s_steps, s_att, s_ops);
        fflush(stdout);
        tot_steps += s_steps;
        tot_att += s_att;
        tot_ops += s_ops;
        tot_traj += n;
    }
    printf("total traj %llu steps %llu attempts %llu ops %llu ops_per_step %.3f\n", tot_traj, tot_steps,
           tot_att, tot_ops, (double)tot_ops / (double)tot_steps);
    return 0;
}
