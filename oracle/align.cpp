// align.cpp -- per-pair alignment driver, Powell refinement and Fourier averaging of the oracle, plus the
// extern "C" functions the tests load through ctypes.
// TEST INFRASTRUCTURE ONLY (the checker); see oracle_core.hpp.
//
// Restates nbfProjectionRotationMetric3D::execute (ProjRot.h:416-1052), getTrueLocalMaximaPositions
// (:1593-1634), eliminateEquivalentAngles (:1637-1702), localMaxima3D3x3 (:67-79) with
// BordStrategyConst(C1, 2, -max) (inc/bs/nbfBordStrategyConst.h), PriorityQueue/greaterPD
// (inc/fm/nbfFastMarching.h:18-40), nbfFourierImageMetric::refine/f (FIM.h:1702-1809, :129-190) with the
// vnl_powell / vnl_bracket_minimum / vnl_brent_minimizer algorithms they call (ITK 4.2.1
// Modules/ThirdParty/VNL/src/vxl/core/vnl/algo/vnl_powell.cxx:60-175, vnl_bracket_minimum.cxx,
// vnl_brent_minimizer.cxx) and nbfWedgedAverageImage3D::getImageMPI/getImage (WAvg.h:217-356).
#include "oracle_core.hpp"

#include <limits>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <queue>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

struct CandCmp {
    bool operator()(const Cand& x, const Cand& y) const { return x.v > y.v; }  // greaterPD: min-heap on [3]
};
using CandQueue = std::priority_queue<Cand, std::vector<Cand>, CandCmp>;

static void eliminate_equivalent(std::vector<Cand>& pos, CandQueue& pq, float tol) {
    const float NEG = -FLT_MAX;
    for (size_t i = 0; i < pos.size(); i++) {
        M4 m1 = m4_zyz(pos[i].a, pos[i].b, pos[i].g);
        for (size_t j = 0; j < pos.size(); j++) {
            if (pos[i].v == NEG) break;
            if (i != j && pos[j].v > NEG) {
                M4 m2 = m4_zyz(pos[j].a, pos[j].b, pos[j].g);
                double err = 0;
                for (int m = 0; m < 16; m++) err += std::fabs(m1.a[m] - m2.a[m]);
                if (err < tol) {
                    if (pos[i].v > pos[j].v) pos[i].v = NEG;
                    else pos[j].v = NEG;
                }
            }
        }
    }
    for (auto& p : pos)
        if (p.v > NEG) pq.push(p);
}

static void restrict_beta(std::vector<double>& C, int B, double rot) {
    const int n = 2 * B;
    double f = std::floor(rot * B / 90.0 - 1.0);
    int b0 = (int)std::max(1.0, std::min(f, (double)(n - 1)));
    for (int a = 0; a < n; a++)
        for (int b = b0; b < n; b++)
            for (int g = 0; g < n; g++) C[((size_t)a * n + b) * n + g] = -(double)FLT_MAX;
}

static void local_maxima(const std::vector<double>& C, int n, std::vector<double>& out) {
    const double border = -(double)FLT_MAX;
    auto at = [&](int a, int b, int g) -> double {
        if (a < 0 || b < 0 || g < 0 || a >= n || b >= n || g >= n) return border;
        return C[((size_t)a * n + b) * n + g];
    };
    out.assign(C.size(), 0.0);
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++)
            for (int g = 0; g < n; g++) {
                double v = at(a, b, g);
                bool ok = true;
                for (int da = -1; da <= 1 && ok; da++)
                    for (int db = -1; db <= 1 && ok; db++)
                        for (int dg = -1; dg <= 1 && ok; dg++) {
                            if (!da && !db && !dg) continue;
                            if (!(v >= at(a + da, b + db, g + dg))) ok = false;
                        }
                out[((size_t)a * n + b) * n + g] = ok ? v : 0.0;
            }
}

static void true_local_maxima(const Plan& P, const std::vector<double>& C, CandQueue& pq, float tol) {
    const int B = P.p.B, n = 2 * B;
    std::vector<Cand> pos;
    double rot = std::max((double)P.p.rot_restrict, 1.0);
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++)
            for (int g = 0; g < n; g++) {
                double v = C[((size_t)a * n + b) * n + g];
                if (std::fabs(v) > 0) {
                    float alpha = (float)(a / (double)B * 180.0);
                    float beta = (float)((2 * b + 1.0) / (4. * B) * 180.0);
                    if ((float)B <= 32) beta = (float)((2 * b + 0.0) / (4. * B) * 180.0);
                    float gamma = (float)(g / (double)B * 180.0);
                    M4 t = m4_zyz(alpha, beta, gamma);
                    double tr = P.p.trans_restrict < 0 ? (double)FLT_MAX : (double)P.p.trans_restrict;
                    if (transform_valid(t, tr, rot)) pos.push_back(Cand{alpha, beta, gamma, (float)(-v)});
                }
            }
    eliminate_equivalent(pos, pq, tol);
}

std::vector<Cand> candidates_of(const Plan& P, std::vector<double> C, float tol) {
    const int n = 2 * P.p.B;
    restrict_beta(C, P.p.B, std::max((double)P.p.rot_restrict, 1.0));
    std::vector<double> L;
    local_maxima(C, n, L);
    CandQueue pq;
    true_local_maxima(P, L, pq, tol);
    std::vector<Cand> out;
    while (!pq.empty()) {
        out.push_back(pq.top());
        pq.pop();
    }
    return out;
}

// --------------------------------------------------------------------------- vnl Powell restatement
namespace {
struct PowellState {
    std::function<double(const double*)> f;
    double end_error = 0;
    long nevals = 0;
    void report(double e) {
        if (nevals == 0) end_error = e;
        if (e < end_error) end_error = e;
        nevals++;
    }
};
const double GOLD = 1.618033988749894848, CGOLD = 0.381966011250105152;

struct Line {
    PowellState* S;
    double x0[3], dx[3];
    double operator()(double lam) {
        double x[3] = {x0[0] + lam * dx[0], x0[1] + lam * dx[1], x0[2] + lam * dx[2]};
        double e = S->f(x);
        S->report(e);
        return e;
    }
};

void bracket_minimum(Line& f, double& a, double& b, double& c) {
    const double EPS = 1e-7, EPSqr = 1e-14;
    if (b == a) b = a + 1.0;
    double fa = f(a), fb = f(b);
    if (fb > fa) { std::swap(a, b); std::swap(fa, fb); }
    c = b + GOLD * (b - a);
    double fc = f(c);
    while (fc < fb) {
        double x1 = a - b, f1 = fa - fb, x2 = c - b, f2 = fc - fb;
        double p = x2 * x2 * f1 - x1 * x1 * f2, q = 2 * (x2 * f1 - x1 * f2);
        if (q >= 0 && q < EPSqr) q = EPSqr;
        else if (q < 0 && q + EPSqr > 0) q = -1.0 * EPSqr;
        double du = p / q;
        double tol = EPS * (1.0 + std::max(std::fabs(b), std::fabs(c)));
        if (du >= 0 && du < tol) du = tol;
        else if (du < 0 && du + tol > 0) du = -1.0 * tol;
        double u = b + du;
        if ((u - c) < tol && (u - c) >= 0) u += tol;
        else if ((c - u) < tol && (c - u) >= 0) u -= tol;
        double u_limit = b + 100 * (c - b);
        double fu = 0.0;
        if ((u - b) * (c - u) > 0.0) {
            fu = f(u);
            if (fu < fc) {
                a = b; fa = fb; b = u; fb = fu;
                if (a > c) { std::swap(a, c); std::swap(fa, fc); }
                return;
            } else if (fu > fb) {
                c = u; fc = fu;
                if (a > c) { std::swap(a, c); std::swap(fa, fc); }
                return;
            }
            u = c + GOLD * (c - b);
            fu = f(u);
        } else if ((u - c) * (u_limit - u) > 0.0) {
            fu = f(u);
            if (fu > fc) {
                a = b; fa = fb; b = c; fb = fc; c = u; fc = fu;
                if (a > c) { std::swap(a, c); std::swap(fa, fc); }
                return;
            }
        } else if ((u_limit - c) * (u - u_limit) >= 0) {
            u = u_limit;
            fu = f(u);
        } else {
            u = c + GOLD * (c - b);
            fu = f(u);
        }
        a = b; b = c; c = u;
        fa = fb; fb = fc; fc = fu;
    }
    if (a > c) { std::swap(a, c); std::swap(fa, fc); }
}

double brent_given_bounds(Line& f, double ax, double bx, double cx, double xtol, double& fmin) {
    const double EPS = 1e-8;
    double fb = f(bx);
    double x = bx, w = x, v = w, u;
    double fx = fb, fw = fx, fv = fw, fu;
    double m = 0.5 * (ax + cx), tol = EPS * std::fabs(x) + xtol, tol2 = 2 * tol;
    double d = 0.0, e = 0.0;
    while (std::fabs(x - m) > (tol2 - 0.5 * (cx - ax))) {
        double p = 0.0, q = 0.0, r = 0.0;
        if (std::fabs(e) > tol) {
            r = (x - w) * (fx - fv);
            q = (x - v) * (fx - fw);
            p = (x - v) * q - (x - w) * r;
            q = 2 * (q - r);
            if (q > 0) p *= -1.0; else q *= -1.0;
            r = e; e = d;
        }
        if (std::fabs(p) < std::fabs(0.5 * q * r) && p > (q * (ax - x)) && p < (q * (cx - x))) {
            d = p / q;
            u = x + d;
            if (u - ax < tol2 || cx - u < tol2) d = (x < m ? tol : -tol);
        } else {
            e = (x < m ? cx : ax) - x;
            d = CGOLD * e;
        }
        if (std::fabs(d) >= tol) u = x + d;
        else u = (d > 0) ? x + tol : x - tol;
        fu = f(u);
        if (fu <= fx) {
            if (u < x) cx = x; else ax = x;
            v = w; fv = fw; w = x; fw = fx; x = u; fx = fu;
        } else {
            if (u < x) ax = u; else cx = u;
            if (fu <= fw || w == x) { v = w; fv = fw; w = u; fw = fu; }
            else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
        }
        m = 0.5 * (ax + cx);
        tol = EPS * std::fabs(x) + xtol;
        tol2 = 2 * tol;
    }
    fmin = fx;
    return x;
}

// vnl_powell::minimize with maxfev = 1, initial_step, linmin_xtol, f_tolerance
void powell_minimize(PowellState& S, double p[3], double initial_step, double linmin_xtol, double ftol) {
    const int n = 3;
    double xi[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double ptt[3], xit[3], pt[3];
    double fret = S.f(p);
    S.report(fret);
    for (int j = 0; j < n; j++) pt[j] = p[j];
    int num_iterations = 0;
    const int maxfev = 1;
    while (num_iterations < maxfev) {
        double fp = fret;
        int ibig = 0;
        double del = 0.0;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) xit[j] = xi[j][i];
            double fptt = fret;
            Line L{&S, {p[0], p[1], p[2]}, {xit[0], xit[1], xit[2]}};
            double ax = 0.0, xx = initial_step, bx;
            bracket_minimum(L, ax, xx, bx);
            double fm;
            xx = brent_given_bounds(L, ax, xx, bx, linmin_xtol, fm);
            fret = fm;
            for (int j = 0; j < n; j++) p[j] = L.x0[j] + xx * L.dx[j];
            if (std::fabs(fptt - fret) > del) { del = std::fabs(fptt - fret); ibig = i; }
        }
        if (2.0 * std::fabs(fp - fret) <= ftol * (std::fabs(fp) + std::fabs(fret))) return;
        if (num_iterations == maxfev) return;
        for (int j = 0; j < n; j++) {
            ptt[j] = 2.0 * p[j] - pt[j];
            xit[j] = p[j] - pt[j];
            pt[j] = p[j];
        }
        double fptt = S.f(ptt);
        S.report(fret);
        if (fptt < fp) {
            double t = 2.0 * (fp - 2.0 * fret + fptt) * (fp - fret - del) * (fp - fret - del) - del * (fp - fptt) * (fp - fptt);
            if (t < 0.0) {
                Line L{&S, {p[0], p[1], p[2]}, {xit[0], xit[1], xit[2]}};
                double ax = 0.0, xx = 1.0, bx;
                bracket_minimum(L, ax, xx, bx);
                double fm;
                xx = brent_given_bounds(L, ax, xx, bx, linmin_xtol, fm);
                fret = fm;
                for (int j = 0; j < n; j++) p[j] = L.x0[j] + xx * L.dx[j];
                for (int j = 0; j < n; j++) {
                    xi[j][ibig] = xi[j][n - 1];
                    xi[j][n - 1] = xit[j];
                }
            }
        }
        num_iterations++;
    }
}
}  // namespace

static M4 trial(const M4& seed, const double* x) {
    M4 t = seed;
    pre(t, m4_rx(x[0]));
    pre(t, m4_ry(x[1]));
    pre(t, m4_rz(x[2]));
    return t;
}

void powell_refine(Plan& P, const float* part, const ettk_particle& d, M4& t, bool run) {
    const double rot = std::max((double)P.p.rot_restrict, 1.0);
    const double tr = P.p.trans_restrict < 0 ? (double)FLT_MAX : (double)P.p.trans_restrict;
    M4 seed = m4_strip_translation(t);
    PowellState S;
    S.f = [&](const double* x) -> double {
        M4 tt = trial(seed, x);
        if (!transform_valid(tt, tr, rot)) return (double)FLT_MAX;
        float peak = ccc_eval(P, part, d, tt);
        if (!transform_valid(tt, tr, rot)) return (double)FLT_MAX;
        return (double)peak;
    };
    double x[3] = {0, 0, 0};
    if (run) powell_minimize(S, x, 5.0, 1e-2, 1e-1);
    if (S.end_error != (double)FLT_MAX) {
        M4 vt = trial(seed, x);
        ccc_eval(P, part, d, vt);
        t = vt;
    } else {
        ccc_eval(P, part, d, t);
    }
}

// --------------------------------------------------------------------------- per-pair driver
static float ccc_tracked(Plan& P, const float* part, const ettk_particle& d, M4& t) {
    return ccc_eval(P, part, d, t);
}

void align_pair(Plan& P, const float* part, const ettk_particle& d, ettk_result& out) {
    const int B = P.p.B, n = 2 * B, nn = n * n;
    const double rot = std::max((double)P.p.rot_restrict, 1.0);
    const double tr = P.p.trans_restrict < 0 ? (double)FLT_MAX : (double)P.p.trans_restrict;
    std::memset(&out, 0, sizeof(out));
    int ccc0 = P.ccc_count;
    // projections (ProjRot.h:441-470)
    std::vector<float> Pp1 = P.Pp1, Pp2;
    std::vector<double> img2;
    reslice_image(part, P.N, d, nullptr, img2);
    project_image(P, img2, Pp2);
    float M = -FLT_MAX;
    for (float v : Pp1) M = std::max(M, v);
    for (float v : Pp2) M = std::max(M, v);
    for (auto& v : Pp1) v = v / M;
    for (auto& v : Pp2) v = v / M;
    std::vector<float> W1(nn), W2(nn);
    bool uw1 = false, uw2 = false;
    if (P.p.missing_wedge) {
        W1 = P.ref_wedge_kind == 0 ? spherical_wedge(B, P.refd) : P.W1s_acc;  // average: WAvg.h:670-690
        W2 = spherical_wedge(B, d);
        condition_spherical_window(W1, B);
        condition_spherical_window(W2, B);
        double s1 = 0, s2 = 0;
        for (int i = 0; i < nn; i++) { s1 += W1[i]; s2 += W2[i]; }
        uw1 = s1 <= 0.95f * nn;
        uw2 = s2 <= 0.95f * nn;
    }
    out.use_window = (uw1 ? 1 : 0) | (uw2 ? 2 : 0);
    std::vector<double> F1(nn), F2(nn), C1, C2;
    if (uw1 && uw2) {
        for (int i = 0; i < nn; i++) { F1[i] = Pp1[i] * Pp1[i] * W1[i]; F2[i] = W2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C1);
        for (int i = 0; i < nn; i++) { F1[i] = W1[i]; F2[i] = Pp2[i] * Pp2[i] * W2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C2);
        for (size_t o = 0; o < C1.size(); o++) C1[o] = std::sqrt(C1[o]) * std::sqrt(C2[o]);
        for (int i = 0; i < nn; i++) { F1[i] = Pp1[i] * W1[i]; F2[i] = Pp2[i] * W2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C2);
        for (size_t o = 0; o < C1.size(); o++) C1[o] = 2.0 * (1.0 - C2[o] / C1[o]);
        for (int i = 0; i < nn; i++) { F1[i] = W1[i]; F2[i] = W2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C2);
        for (size_t o = 0; o < C1.size(); o++) C1[o] = -C1[o] / C2[o];
    } else if (uw1) {
        for (int i = 0; i < nn; i++) { F1[i] = W1[i]; F2[i] = Pp2[i] * Pp2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C2);
        for (int i = 0; i < nn; i++) { F1[i] = W1[i] * Pp1[i]; F2[i] = Pp2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C1);
        for (size_t o = 0; o < C1.size(); o++) C1[o] = C1[o] / std::sqrt(C2[o]);
    } else if (uw2) {
        for (int i = 0; i < nn; i++) { F1[i] = Pp1[i] * Pp1[i]; F2[i] = W2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C2);
        for (int i = 0; i < nn; i++) { F1[i] = Pp1[i]; F2[i] = W2[i] * Pp2[i]; }
        so3_correlate_scaled(B, F1.data(), F2.data(), C1);
        for (size_t o = 0; o < C1.size(); o++) C1[o] = C1[o] / std::sqrt(C2[o]);
    } else {
        for (int i = 0; i < nn; i++) { F1[i] = Pp1[i]; F2[i] = Pp2[i]; }
        so3_correlate(B, F1.data(), F2.data(), C1);
    }
    restrict_beta(C1, B, rot);
    std::vector<double> LM;
    local_maxima(C1, n, LM);
    CandQueue pq1, pq2;
    true_local_maxima(P, LM, pq1, 0.25f);
    if (uw1 || uw2) {
        for (int i = 0; i < nn; i++) { F1[i] = Pp1[i]; F2[i] = Pp2[i]; }
        so3_correlate(B, F1.data(), F2.data(), C2);
        restrict_beta(C2, B, rot);
        local_maxima(C2, n, LM);
        true_local_maxima(P, LM, pq2, 0.25f);
    }
    // direct + mirror CCC at the candidate peaks (ProjRot.h:678-776)
    std::vector<Cand> combined;
    int nq = (uw1 || uw2) ? 2 : 1;
    for (int q = 0; q < nq; q++) {
        CandQueue& pq = (q == 0) ? pq1 : pq2;
        int peaks = P.p.n_peaks;
        if (q == 1) peaks = (int)std::max(1.0, std::floor(P.p.n_peaks / 3.0));
        int index = 0;
        while (index < peaks && !pq.empty()) {
            Cand p = pq.top();
            if (p.v == FLT_MAX) break;
            M4 t1 = m4_zyz(p.a, p.b, p.g);
            M4 td = t1;
            float D = ccc_tracked(P, part, d, td);
            if (transform_valid(td, tr, rot)) {
                M4 mir = m4_mul(m4_scale(-1, -1, -1), t1);
                float Dm = ccc_tracked(P, part, d, mir);
                if (Dm >= D) index++;
                combined.push_back(Cand{p.a, p.b, p.g, D});
            }
            pq.pop();
        }
    }
    if (combined.empty()) combined.push_back(Cand{0, 0, 0, 1.0f});
    CandQueue pqc;
    eliminate_equivalent(combined, pqc, 2.5f);
    float best = FLT_MAX;
    M4 bestT = m4_ident();
    int index = 0;
    while (!pqc.empty() && index < P.p.n_candidates && index < ETTK_MAX_CANDIDATES) {
        Cand p = pqc.top();
        M4 t = m4_zyz(p.a, p.b, p.g);
        powell_refine(P, part, d, t, P.p.refine != 0);
        float dist = P.last_peak;
        if (dist < best) {
            best = dist;
            bestT = t;
        }
        out.cand_distance[index] = dist;
        std::memcpy(out.cand_matrix[index], t.a, sizeof(t.a));
        out.so3_angles[index][0] = p.a;
        out.so3_angles[index][1] = p.b;
        out.so3_angles[index][2] = p.g;
        pqc.pop();
        index++;
    }
    out.n_cand = index;
    out.distance = best;
    std::memcpy(out.matrix, bestT.a, sizeof(bestT.a));
    out.n_ccc = P.ccc_count - ccc0;
}

}  // namespace orc

// =========================================================================== extern "C"
using namespace orc;

extern "C" {

void* orc_plan_create(const ettk_params* p) {
    Plan* P = new Plan();
    plan_init(*P, *p);
    return P;
}
void orc_plan_destroy(void* h) { delete (Plan*)h; }

void orc_set_reference(void* h, const float* ref, const ettk_particle* d) {
    Plan& P = *(Plan*)h;
    plan_set_reference(P, ref, *d);
    project_image(P, P.ref, P.Pp1);  // input1 = the reference image as it is (ProjRot.h:445-453)
}

float orc_ccc(void* h, const float* part, const ettk_particle* d, double* t_inout) {
    Plan& P = *(Plan*)h;
    M4 t = m4_from(t_inout);
    float v = ccc_eval(P, part, *d, t);
    std::memcpy(t_inout, t.a, sizeof(t.a));
    return v;
}

void orc_project(void* h, const float* vol, const ettk_particle* d, int is_reference, float* out) {
    Plan& P = *(Plan*)h;
    std::vector<double> img;
    if (is_reference) to_internal(vol, P.N, img);
    else reslice_image(vol, P.N, *d, nullptr, img);
    std::vector<float> Pp;
    project_image(P, img, Pp);
    std::memcpy(out, Pp.data(), Pp.size() * sizeof(float));
}

void orc_get_pp1(void* h, float* out) {
    Plan& P = *(Plan*)h;
    std::memcpy(out, P.Pp1.data(), P.Pp1.size() * sizeof(float));
}

void orc_align(void* h, const float* part, const ettk_particle* d, ettk_result* out) {
    align_pair(*(Plan*)h, part, *d, *out);
}

// OpenMP over particles, one plan copy per thread (one alignment job per MPI rank in the reference)
void orc_align_many(void* h, int n, const float* parts, const ettk_particle* d, ettk_result* out, int nthreads) {
    Plan& P0 = *(Plan*)h;
    const size_t n3 = (size_t)P0.N * P0.N * P0.N;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        Plan P = P0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < n; i++) align_pair(P, parts + (size_t)i * n3, d[i], out[i]);
    }
}

double orc_refine(void* hp, const float* part, const ettk_particle* d, const double* seed, int n_iter, double* pose_out);

// The batched + adjoint form of the library in FP64, OpenMP over particles: the search and candidate stage of
// the reference without Powell (refinementOff: one CCC at each selected candidate, FIM.h:1758-1766), then
// n_iter steps of the continuous ascent (orc_refine) from the best candidate's transform, then one CCC at the
// refined pose for the reported distance (the refined pose is kept).  scores[i] = final continuous score.
void orc_align_adjoint_many(void* h, int n, const float* parts, const ettk_particle* d, int n_iter, ettk_result* out,
                            double* scores, int nthreads) {
    Plan& P0 = *(Plan*)h;
    const size_t n3 = (size_t)P0.N * P0.N * P0.N;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        Plan P = P0;
        P.p.refine = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < n; i++) {
            const float* part = parts + (size_t)i * n3;
            align_pair(P, part, d[i], out[i]);
            double pose[16];
            scores[i] = orc_refine(&P, part, &d[i], out[i].matrix, n_iter, pose);
            M4 t = m4_from(pose);
            out[i].distance = ccc_eval(P, part, d[i], t);
            out[i].cand_distance[0] = out[i].distance;
            std::memcpy(out[i].matrix, pose, sizeof(pose));
            std::memcpy(out[i].cand_matrix[0], pose, sizeof(pose));
            out[i].n_ccc += 1;
        }
    }
}

void orc_s2_forward(int B, const double* f, double* re, double* im) {
    std::vector<cd> c;
    s2_forward(B, f, c);
    for (size_t i = 0; i < c.size(); i++) { re[i] = c[i].real(); im[i] = c[i].imag(); }
}
void orc_s2_inverse(int B, const double* re, const double* im, double* f) {
    std::vector<cd> c((size_t)B * B);
    for (size_t i = 0; i < c.size(); i++) c[i] = cd(re[i], im[i]);
    s2_inverse(B, c, f);
}
void orc_so3_correlate(int B, const double* sig, const double* pat, double* C) {
    std::vector<double> c;
    so3_correlate(B, sig, pat, c);
    std::memcpy(C, c.data(), c.size() * sizeof(double));
}
void orc_so3_correlate_scaled(int B, const double* sig, const double* pat, double* C) {
    std::vector<double> c;
    so3_correlate_scaled(B, sig, pat, c);
    std::memcpy(C, c.data(), c.size() * sizeof(double));
}
void orc_s2_rotate(int B, const double* in, double* out, double a, double b, double g) { s2_rotate(B, in, out, a, b, g); }

// harmonics(A, t, R) (ProjRot.h:1348-1454): Euler angles of the rotation part of t in Pixel (float) arithmetic
// (:1369-1448), rotateFct by (-psi, -theta, -phi) of the image as REAL (double), result cast to Pixel
void orc_harmonics_t(int B, const float* A, const double* t, float* R) {
    const double m20 = t[8], m21 = t[9], m12 = t[6], m02 = t[2], m22 = t[10];
    const double fmin = (double)std::numeric_limits<float>::min();
    float psi, phi, theta;
    if (std::fabs(m20) < fmin) psi = std::fabs(m21) < fmin ? 0.0f : (m21 < -fmin ? (float)(M_PI / 2.0) : (float)(-M_PI / 2.0));
    else psi = (float)std::atan(-m21 / m20);
    if (std::fabs(m02) < fmin) phi = std::fabs(m12) < fmin ? 0.0f : (m12 > fmin ? (float)(M_PI / 2.0) : (float)(-M_PI / 2.0));
    else phi = (float)std::atan(m12 / m02);
    theta = (float)std::acos(m22);
    float psi_1 = psi, psi_2 = psi > 0 ? (float)(psi - M_PI) : (float)(psi + M_PI);
    float phi_1 = phi, phi_2 = phi > 0 ? (float)(phi - M_PI) : (float)(phi + M_PI);
    psi = (m21 * psi_1 > 0) ? psi_1 : psi_2;   // theta > 0
    phi = (m12 * phi_1 > 0) ? phi_1 : phi_2;
    if (std::cos(psi) * m20 > 0) {              // theta < 0
        theta = -theta;
        psi = (m21 * psi_1 < 0) ? psi_1 : psi_2;
        phi = (m12 * phi_1 < 0) ? phi_1 : phi_2;
    }
    const int n = 2 * B;
    std::vector<double> a((size_t)n * n), r((size_t)n * n);
    for (size_t i = 0; i < a.size(); i++) a[i] = (double)A[i];
    s2_rotate(B, a.data(), r.data(), -(double)psi, -(double)theta, -(double)phi);
    for (size_t i = 0; i < r.size(); i++) R[i] = (float)r[i];
}
void orc_wigner_d(int B, double beta, double* d) {
    std::vector<double> v;
    wigner_d_all(B, beta, v);
    std::memcpy(d, v.data(), v.size() * sizeof(double));
}
void orc_r2c3(const double* in, double* out_interleaved, int n0, int n1, int n2) {
    r2c3(in, reinterpret_cast<cd*>(out_interleaved), n0, n1, n2);
}
void orc_c2r3(const double* in_interleaved, double* out, int n0, int n1, int n2) {
    c2r3(reinterpret_cast<const cd*>(in_interleaved), out, n0, n1, n2);
}
void orc_reslice(const float* vol, int N, const ettk_particle* d, const double* t, double* out) {
    std::vector<double> o;
    M4 T;
    if (t) T = m4_from(t);
    reslice_image(vol, N, *d, t ? &T : nullptr, o);
    std::memcpy(out, o.data(), o.size() * sizeof(double));
}
void orc_wedge_half(int N, const ettk_particle* d, const double* t, float* out) {
    M4 T;
    if (t) T = m4_from(t);
    std::vector<float> w = particle_wedge_half(N, *d, t ? &T : nullptr);
    std::memcpy(out, w.data(), w.size() * sizeof(float));
}
void orc_spherical_wedge(int B, const ettk_particle* d, int condition, float* out) {
    std::vector<float> w = spherical_wedge(B, *d);
    if (condition) condition_spherical_window(w, B);
    std::memcpy(out, w.data(), w.size() * sizeof(float));
}
// accumulated wedge of an average (WAvg.h:49-70, 496-538, 614-667); outputs may be null
void orc_avg_wedge(int N, int B, int n, const ettk_particle* d, const float* w, const double* mats, float* accum,
                   float* bin, float* wedge, float* sph, int* effective) {
    AvgWedge aw = avg_wedge(N, B, n, d, w, mats);
    if (accum) std::memcpy(accum, aw.accum.data(), aw.accum.size() * sizeof(float));
    if (bin) std::memcpy(bin, aw.bin.data(), aw.bin.size() * sizeof(float));
    if (wedge) std::memcpy(wedge, aw.wedge.data(), aw.wedge.size() * sizeof(float));
    if (sph)
        for (size_t o = 0; o < aw.sph.size(); o++) sph[o] = aw.effective ? aw.sph[o] : 1.0f;
    if (effective) *effective = aw.effective ? 1 : 0;
}
void orc_set_reference_average(void* h, const float* ref, int n, const ettk_particle* d, const float* w,
                               const double* mats) {
    Plan& P = *(Plan*)h;
    AvgWedge aw = avg_wedge(P.N, P.p.B, n, d, w, mats);
    plan_set_reference_average(P, ref, aw);
    project_image(P, P.ref, P.Pp1);
}

float orc_wedge_overlap(void* h, const ettk_particle* d, const double* t) {
    return wedge_overlap(*(Plan*)h, *d, m4_from(t));
}

void orc_masks(void* h, float* maskfft, float* refmask, float* filter) {
    Plan& P = *(Plan*)h;
    std::memcpy(maskfft, P.maskfft.data(), P.maskfft.size() * sizeof(float));
    std::memcpy(refmask, P.refmask.data(), P.refmask.size() * sizeof(float));
    std::memcpy(filter, P.filter.data(), P.filter.size() * sizeof(float));
}
int orc_ccc_count(void* h) { return ((Plan*)h)->ccc_count; }

// executeFourierNewHalf with its search window: T (ext^3 doubles, [i][j][k], before the spherical mask) and
// info = {k1_m, ext}; returns correlationPeak, t_inout <- t * Translate(-shift)
float orc_ccc_trace(void* h, const float* part, const ettk_particle* d, double* t_inout, double* T, double* info) {
    Plan& P = *(Plan*)h;
    M4 t = m4_from(t_inout);
    CccTrace tr;
    float v = ccc_eval(P, part, *d, t, &tr);
    std::memcpy(t_inout, t.a, sizeof(t.a));
    if (T) std::memcpy(T, tr.T.data(), tr.T.size() * sizeof(double));
    info[0] = tr.k1_m;
    info[1] = tr.ext;
    return v;
}

// sub-volumes cut from a volume (x fastest, nx x ny x nz) at positions pos[3 i]: out[i] is N^3 doubles, x fastest
void orc_extract(const float* vol, int nx, int ny, int nz, int n, const double* pos, const ettk_particle* d, int N,
                 double* out) {
    const size_t n3 = (size_t)N * N * N;
    for (int i = 0; i < n; i++) {
        std::vector<double> A;
        extract_image(vol, nx, ny, nz, pos + 3 * i, N, d[i], nullptr, A);
        for (int x = 0; x < N; x++)
            for (int y = 0; y < N; y++)
                for (int z = 0; z < N; z++) out[i * n3 + ((size_t)z * N + y) * N + x] = A[((size_t)x * N + y) * N + z];
    }
}

// the surviving candidates of one correlation volume, in priority order: out[4 i] = (alpha, beta, gamma, -C)
int orc_candidates(void* h, const double* C, float tol, float* out, int max_out) {
    Plan& P = *(Plan*)h;
    const size_t n = (size_t)2 * P.p.B;
    std::vector<double> c(C, C + n * n * n);
    std::vector<Cand> v = candidates_of(P, c, tol);
    const int m = std::min((int)v.size(), max_out);
    for (int i = 0; i < m; i++) {
        out[4 * i] = v[i].a;
        out[4 * i + 1] = v[i].b;
        out[4 * i + 2] = v[i].g;
        out[4 * i + 3] = v[i].v;
    }
    return (int)v.size();
}

// nbfImageMetric::isTransformValid (nbfImageMetric.h:416-459); trans_restrict < 0 is unrestricted
int orc_transform_valid(const double* t, double trans_restrict, double rot_restrict) {
    return transform_valid(m4_from(t), trans_restrict < 0 ? (double)FLT_MAX : trans_restrict, std::max(rot_restrict, 1.0)) ? 1 : 0;
}

// Amplitude-weighted Fourier average (WAvg.h:261-356, WAvg.h:217-258); out is N^3 doubles, x fastest.
void orc_average(void* h, int n, const float* parts, const ettk_particle* d, const float* weights,
                 const double* mats, double* out) {
    Plan& P = *(Plan*)h;
    const int N = P.N, H = P.H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    std::vector<double> re(nh, 0.0), im(nh, 0.0), acc(nh, 0.0), T, A(n3);
    std::vector<cd> F(nh);
    for (int i = 0; i < n; i++) {
        if (!(std::fabs(weights[i]) > 0)) continue;
        ettk_particle di = d[i];
        for (int k = 0; k < 16; k++) di.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;  // setTransform(NULL)
        M4 store = m4_from(mats + (size_t)i * 16);
        reslice_image(parts + (size_t)i * n3, N, di, &store, T);
        for (size_t o = 0; o < n3; o++) A[o] = T[o] * P.shift[o];
        r2c3(A.data(), F.data(), N, N, N);
        double w = weights[i];
        for (size_t o = 0; o < nh; o++) {
            double Bm = std::sqrt(std::real(F[o] * std::conj(F[o])));
            acc[o] += std::fabs(w) * Bm;
            re[o] += w * Bm * F[o].real();
            im[o] += w * Bm * F[o].imag();
        }
    }
    for (size_t o = 0; o < nh; o++) F[o] = acc[o] > 0 ? cd(re[o] / acc[o], im[o] / acc[o]) : cd(0, 0);
    F[((size_t)(N / 2) * N + N / 2) * H + (H - 1)] = 0.0;
    c2r3(F.data(), A.data(), N, N, N);
    double ss = 0;
    for (size_t o = 0; o < n3; o++) {
        A[o] = A[o] / (double)n3 * P.shift[o];
        ss += A[o] * A[o];
    }
    float var_inside = (float)std::sqrt(ss / (double)n3);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) out[((size_t)k * N + j) * N + i] = A[((size_t)i * N + j) * N + k] / var_inside;
}

// Variance map (WAvg.h:440-478): var += |w| (T - avg)^2 / sum|w| over the entries with |w| > 0, T the resliced
// particle (getImage(image, store, false): valid voxels minus their mean, 0 outside); avg and out are x fastest.
void orc_variance(void* h, int n, const float* parts, const ettk_particle* d, const float* weights,
                  const double* mats, const double* avg, double* out) {
    Plan& P = *(Plan*)h;
    const int N = P.N;
    const size_t n3 = (size_t)N * N * N;
    double wsum = 0;
    for (int i = 0; i < n; i++) wsum += std::fabs((double)weights[i]);
    std::vector<double> var(n3, 0.0), T;
    for (int i = 0; i < n; i++) {
        if (!(std::fabs(weights[i]) > 0)) continue;
        ettk_particle di = d[i];
        for (int k = 0; k < 16; k++) di.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;  // setTransform(NULL)
        M4 store = m4_from(mats + (size_t)i * 16);
        reslice_image(parts + (size_t)i * n3, N, di, &store, T);
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++)
                for (int c = 0; c < N; c++) {
                    const size_t o = ((size_t)a * N + b) * N + c;  // A(i,j,k), i = x
                    const double dv = T[o] - avg[((size_t)c * N + b) * N + a];
                    var[o] += std::fabs((double)weights[i]) * (dv * dv) / wsum;
                }
    }
    for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++)
            for (int c = 0; c < N; c++) out[((size_t)c * N + b) * N + a] = var[((size_t)a * N + b) * N + c];
}

// ---------------------------------------------------------------------------- checks of the restated pieces
// Powell (maxfev = 1, the library's settings) on f(x) = sum_i a_i (x_i - c_i)^2 + k (x_0 - c_0)(x_1 - c_1);
// returns the number of function evaluations, writes the minimiser and its value.
int orc_powell_quadratic(const double* a, const double* c, double k, const double* x0, double step, double xtol,
                         double ftol, double* xmin, double* fmin) {
    PowellState S;
    S.f = [&](const double* x) -> double {
        double v = k * (x[0] - c[0]) * (x[1] - c[1]);
        for (int i = 0; i < 3; i++) v += a[i] * (x[i] - c[i]) * (x[i] - c[i]);
        return v;
    };
    double p[3] = {x0[0], x0[1], x0[2]};
    powell_minimize(S, p, step, xtol, ftol);
    for (int i = 0; i < 3; i++) xmin[i] = p[i];
    *fmin = S.end_error;
    return (int)S.nevals;
}
// The same sweep on any function of three variables supplied by the caller (a test's closed forms).
int orc_powell_fn(double (*fn)(const double*), const double* x0, double step, double xtol, double ftol,
                  double* xmin, double* fmin) {
    PowellState S;
    S.f = [fn](const double* x) -> double { return fn(x); };
    double p[3] = {x0[0], x0[1], x0[2]};
    powell_minimize(S, p, step, xtol, ftol);
    for (int i = 0; i < 3; i++) xmin[i] = p[i];
    *fmin = S.end_error;
    return (int)S.nevals;
}
// 26-neighbour local maxima of an n^3 volume with a -FLT_MAX border (localMaxima3D3x3 + BordStrategyConst)
void orc_local_maxima(const double* C, int n, double* out) {
    std::vector<double> c(C, C + (size_t)n * n * n), o;
    local_maxima(c, n, o);
    std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// getDistances worker semantics by metric type (CorrMetric.h:1911-1990): 0 search + refine, 3 search only,
// 1 refine from the stored transform (FIM.h:1812-1840), 2 one CCC at the identity (Core.h:35-66)
void orc_distance(void* h, const float* part, const ettk_particle* d, int type, ettk_result* out) {
    Plan& P = *(Plan*)h;
    if (type == 0 || type == 3) {
        const int saved = P.p.refine;
        P.p.refine = type == 0 ? 1 : 0;
        align_pair(P, part, *d, *out);
        P.p.refine = saved;
        return;
    }
    std::memset(out, 0, sizeof(*out));
    const double rot = std::max((double)P.p.rot_restrict, 1.0);
    const double tr = P.p.trans_restrict < 0 ? (double)FLT_MAX : (double)P.p.trans_restrict;
    const int c0 = P.ccc_count;
    M4 r;
    float D;
    if (type == 2) {
        M4 t = m4_ident();
        D = ccc_eval(P, part, *d, t);
        r = transform_valid(t, tr, rot) ? t : m4_ident();
    } else {
        ettk_particle dz = *d;
        for (int k = 0; k < 16; k++) dz.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;
        r = m4_from(d->matrix);
        powell_refine(P, part, dz, r, true);
        D = P.last_peak;
    }
    out->distance = D;
    out->cand_distance[0] = D;
    std::memcpy(out->matrix, r.a, sizeof(r.a));
    std::memcpy(out->cand_matrix[0], r.a, sizeof(r.a));
    out->n_cand = 1;
    out->n_ccc = P.ccc_count - c0;
}

}  // extern "C"
