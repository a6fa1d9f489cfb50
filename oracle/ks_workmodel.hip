// Work model of the device kernels (test infrastructure, host code only): the product's own per-sample and
// per-pair templates (kerrsplat_amd/csrc) instantiated on the host with a counting scalar, so that the counts are
// the operations of the algorithm the kernels run. +, -, *, / and sqrt count 1, each transcendental (exp, log,
// pow, sin, cos, atan2, acos, cbrt, expm1, log1p, tanh, sinh, cosh) counts 1 and is also tallied as special; a
// sincos counts 2. The splat constants come from the library's own host precompute (host_splat_constants).
// oracle/work_model.py feeds it the kept (sample, splat) pairs of a workload and turns the counts into flops per
// ray-sample for bench.py's roofline. Nothing here runs on a GPU or is timed.
#include <cstdio>
#include <vector>
#include "../kerrsplat_amd/csrc/ks_dsample.hpp"
#include "../kerrsplat_amd/csrc/ks_hostsplat.hpp"

namespace ksd {

struct WorkCount { double flops = 0.0, special = 0.0; };
inline WorkCount& work() { static WorkCount w; return w; }
inline void op(double n = 1.0) { work().flops += n; }
inline void op_special(double n = 1.0) { work().flops += n; work().special += n; }

struct Cnt {
    double v;
    Cnt() : v(0.0) {}
    Cnt(double x) : v(x) {}
    Cnt(int x) : v((double)x) {}
};
inline double val(Cnt x) { return x.v; }
#define KSW_BIN(o)                                                               \
    inline Cnt operator o(Cnt a, Cnt b) { op(); return Cnt(a.v o b.v); }         \
    inline Cnt operator o(Cnt a, double b) { op(); return Cnt(a.v o b); }        \
    inline Cnt operator o(double a, Cnt b) { op(); return Cnt(a o b.v); }
KSW_BIN(+)
KSW_BIN(-)
KSW_BIN(*)
KSW_BIN(/)
#undef KSW_BIN
#define KSW_CMP(o)                                                               \
    inline bool operator o(Cnt a, Cnt b) { return a.v o b.v; }                   \
    inline bool operator o(Cnt a, double b) { return a.v o b; }                  \
    inline bool operator o(double a, Cnt b) { return a o b.v; }
KSW_CMP(<)
KSW_CMP(>)
KSW_CMP(<=)
KSW_CMP(>=)
KSW_CMP(==)
KSW_CMP(!=)
#undef KSW_CMP
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { op(); a.v += b.v; return a; }
inline Cnt& operator-=(Cnt& a, Cnt b) { op(); a.v -= b.v; return a; }
inline Cnt& operator*=(Cnt& a, Cnt b) { op(); a.v *= b.v; return a; }
inline Cnt sqrt(Cnt x) { op(); return Cnt(std::sqrt(x.v)); }
#define KSW_SPECIAL(name)                                                        \
    inline Cnt name(Cnt x) { op_special(); return Cnt(std::name(x.v)); }
KSW_SPECIAL(exp)
KSW_SPECIAL(expm1)
KSW_SPECIAL(log)
KSW_SPECIAL(log1p)
KSW_SPECIAL(sin)
KSW_SPECIAL(cos)
KSW_SPECIAL(sinh)
KSW_SPECIAL(cosh)
KSW_SPECIAL(tanh)
KSW_SPECIAL(cbrt)
KSW_SPECIAL(acos)
#undef KSW_SPECIAL
inline void sincos(Cnt x, Cnt* s, Cnt* c) { op_special(2.0); *s = Cnt(std::sin(x.v)); *c = Cnt(std::cos(x.v)); }
inline Cnt atan2(Cnt y, Cnt x) { op_special(); return Cnt(std::atan2(y.v, x.v)); }
inline Cnt pow(Cnt x, Cnt y) { op_special(); return Cnt(std::pow(x.v, y.v)); }
inline Cnt pow(Cnt x, double y) { op_special(); return Cnt(std::pow(x.v, y)); }
inline Cnt pow(double x, Cnt y) { op_special(); return Cnt(std::pow(x, y.v)); }
inline Cnt fabs(Cnt x) { return Cnt(std::fabs(x.v)); }
inline Cnt dmax(Cnt a, Cnt b) { return a.v > b.v ? a : b; }
inline Cnt dmin(Cnt a, Cnt b) { return a.v < b.v ? a : b; }
inline Cnt dmax(Cnt a, double b) { return a.v > b ? a : Cnt(b); }
inline Cnt dmin(Cnt a, double b) { return a.v < b ? a : Cnt(b); }
inline bool isfinite_(Cnt x) { return std::isfinite(x.v); }
inline Cnt pow2_below(Cnt x) { return Cnt(pow2_below(x.v)); }
// the scalar-dependent gates of the step and the pitch angle: those of the FP32 path or of the FP64 path
inline bool& fp32_gates() { static bool g = true; return g; }
inline double sin_floor(Cnt) { return fp32_gates() ? sin_floor(f32(0.0f)) : sin_floor(0.0); }
inline double taylor_gate(Cnt) { return fp32_gates() ? taylor_gate(f32(0.0f)) : taylor_gate(0.0); }

}  // namespace ksd

using namespace ksd;

namespace {
struct Tally { double flops = 0.0, special = 0.0; };
template <class F> Tally count(F&& f) {
    work() = WorkCount();
    f();
    return {work().flops, work().special};
}
}  // namespace

// Counts for `nsmp` active samples of one (frame, frequency) slice of a kappa model whose centres may
// follow knots. Per sample s: its pixel's (alpha, beta, eta, lambda), its (t, r, theta, x, y, z, flags) and its kept
// splats kept[koff[s] .. koff[s + 1]) (ascending splat indices with weight above the cutoff). Output (flops,
// special) pairs, summed over the samples (fp32: the branch gates of the FP32 path, else those of the FP64 path):
//   out[0..1]   the transfer's splat sums (sample_coefficients over the kept splats, the FP64 geometry excluded)
//   out[2..3]   the FP64 geometry (make_geo), once per sample in the transfer (the adjoint reads it back)
//   out[4..5]   the transfer's step (SegMapT::step)
//   out[6..7]   the adjoint's per-sample work (the tail step_apply and sample_adjoint)
//   out[8..9]   the element adjoints (weight_dlog_knot / weight_dlog + element_adjoint_kappa), over every kept pair
//   out[10..11] the loads of the stored geometry: no arithmetic (kept for the layout of the output)
extern "C" int kswm_count_slice(int kind, int nrows, int ns, const double* params, double gamma_max, const double* knots,
                                const double* knot_times, int nknots, double a, double thetao, double t_obs, double nu, double L,
                                int nsmp, const double* pix4, const double* smp7, const int* koff, const int* kept, int fp32,
                                double* out) {
    if (kind != MODEL_KAPPA) return 1;
    fp32_gates() = fp32 != 0;
    const bool with_knots = knots && knot_times && nknots >= 1;
    std::vector<SplatC> h(ns);
    std::vector<SplatAdj> ha(ns);
    for (int i = 0; i < ns; ++i)
        host_splat_constants(kind, nrows, i, params + (size_t)nrows * i, gamma_max, with_knots, knots, nknots, h[i], ha[i]);
    std::vector<SplatCT<Cnt>> sc(ns);
    std::vector<SplatAdjT<Cnt>> sa(ns);
    for (int i = 0; i < ns; ++i) { splat_as(h[i], sc[i]); adj_as(ha[i], sa[i]); }
    std::vector<Cnt> kn(with_knots ? (size_t)3 * nknots * ns : 1), kt(with_knots ? nknots : 1);
    for (size_t q = 0; with_knots && q < kn.size(); ++q) kn[q] = Cnt(knots[q]);
    for (int q = 0; with_knots && q < nknots; ++q) kt[q] = Cnt(knot_times[q]);
    for (int q = 0; q < 12; ++q) out[q] = 0.0;
    const double rp = 1.0 + std::sqrt(1.0 - a * a), rm = 1.0 - std::sqrt(1.0 - a * a);
    std::vector<uint16_t> cand;
    for (int s = 0; s < nsmp; ++s) {
        const double* px = pix4 + 4 * (size_t)s;
        const double* q = smp7 + 7 * (size_t)s;
        const double t = q[0], r = q[1], th = q[2];
        const bool nur = ((int)q[6] & 2) != 0, nuth = ((int)q[6] & 4) != 0;
        const TransCtxT<Cnt> c = make_ctx<Cnt>(a, rp, rm, std::sin(thetao), px[0], px[1], px[2], px[3], sc.data(), ns,
                                               kn.data(), kt.data(), with_knots ? nknots : 0, t_obs, nu, L);
        cand.assign(kept + koff[s], kept + koff[s + 1]);
        Cnt j4[4], a4[4], r3[3];
        bool active = false;
        Tally f = count([&] { active = sample_coefficients<MODEL_KAPPA, Cnt>(c, t, r, th, nur, nuth, Cnt(q[3]), Cnt(q[4]), Cnt(q[5]), j4, a4, r3,
                                                                              cand.data(), (int)cand.size()); });
        out[0] += f.flops; out[1] += f.special;
        GeoT<Cnt> G;
        Tally g = count([&] { make_geo(Cnt(a), Cnt(r), Cnt(th), Cnt(px[2]), Cnt(px[3]), nur, nuth, G); });
        out[2] += g.flops; out[3] += g.special;
        if (!active) continue;
        const double ct = std::cos(th);
        const Cnt D(L / (nu / NU0) * (r * r + a * a * ct * ct) * 1e-3);
        Tally st = count([&] { SegMapT<Cnt> sm; sm.init(); sm.step(j4, a4, r3, D); });
        out[4] += st.flops; out[5] += st.special;
        Cnt w[4] = {Cnt(1.0), Cnt(0.1), Cnt(-0.2), Cnt(0.05)}, R[4] = {Cnt(1e-3), Cnt(1e-5), Cnt(-2e-5), Cnt(1e-6)};
        Tally ad = count([&] {
            Cnt Tn[4], Otw[4], jb[4], ab[4], rb[3];
            step_apply(j4, a4, r3, D, R, Tn);
            sample_adjoint(j4, a4, r3, D, w, R, Otw, jb, ab, rb);
        });
        out[6] += ad.flops; out[7] += ad.special;
        const Cnt bars[BAR_FIELDS] = {Cnt(0.3), Cnt(-0.1), Cnt(0.2), Cnt(0.05), Cnt(-0.4), Cnt(0.1), Cnt(0.07), Cnt(-0.02), Cnt(0.3), Cnt(-0.3), Cnt(0.01)};
        for (int ci = koff[s]; ci < koff[s + 1]; ++ci) {
            const int i = kept[ci];
            Tally e = count([&] {
                Cnt dl[13], grad[ADJ_ROWS];
                for (int r_ = 0; r_ < ADJ_ROWS; ++r_) grad[r_] = Cnt(0.0);
                const Cnt tt(t_obs - t), xx(q[3]), yy(q[4]), zz(q[5]);
                const Cnt Gw = sc[i].nk > 0 ? weight_dlog_knot(sc[i], sa[i], kn.data(), kt.data(), tt, xx, yy, zz, dl)
                                            : weight_dlog(sc[i], sa[i], tt, xx, yy, zz, dl);
                if (Gw > 0.0)
                    element_adjoint_kappa(sc[i], sa[i], G, Cnt(a), Cnt(px[0]), Cnt(px[1]), Cnt(std::sin(thetao)), Cnt(nu), bars, Gw, dl, grad);
            });
            out[8] += e.flops; out[9] += e.special;
        }
    }
    return 0;
}
