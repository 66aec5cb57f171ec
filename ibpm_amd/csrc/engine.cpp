// engine.cpp -- host orchestration of the projection step on one GPU and the
// C ABI declared in include/ibpm_gpu.h.
//
// A batch holds independent flows (uneven grids allowed).  Every kernel launch
// covers all flows of the batch; per-flow descriptors live in device memory
// and are built once at set-up, so a stationary-body timestep is a fixed
// launch sequence that is captured as one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ibpm_gpu.h"
#include "geometry.h"
#include "learn.h"
#include "ops.h"
#include "scratch.h"

namespace ibpmg {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

#define CK(expr)                                                                    \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
            return -1;                                                              \
        }                                                                           \
    } while (0)
#define CKR(expr)                  \
    do {                           \
        int r_ = (expr);           \
        if (r_ != 0) return r_;    \
    } while (0)

// Device allocation owned by a list (freed with the owner).
struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** out, size_t count) {
        void* p = nullptr;
        if (count == 0) count = 1;
        CK(hipMalloc(&p, count * sizeof(T)));
        CK(hipMemset(p, 0, count * sizeof(T)));
        ptrs.push_back(p);
        *out = (T*)p;
        return 0;
    }
};

// Pinned host staging for a flow's per-substep uploads (regulariser lists
// and constraint), so they go as asynchronous copies on the stream.
struct HostStage {
    char* p = nullptr;
    size_t cap = 0, used = 0;
    bool pending = false;          // copies from it may still be in flight
    HostStage() = default;
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    HostStage(HostStage&& o) noexcept : p(o.p), cap(o.cap), used(o.used), pending(o.pending) { o.p = nullptr; }
    ~HostStage() {
        if (p) (void)hipHostFree(p);
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, FftPlan> plans;   // by DST length
    std::map<int, std::pair<__bf16*, __bf16*>> sines;   // split sine matrices of the GEMM-DST, by N
    DevMem mem;
};

// Round to nearest even of an FP32 value to bf16 (the device conversion's rule).
static __bf16 host_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    const uint16_t h = (uint16_t)(u >> 16);
    __bf16 b;
    std::memcpy(&b, &h, 2);
    return b;
}

static float host_float(__bf16 b) {
    uint16_t h;
    std::memcpy(&h, &b, 2);
    const uint32_t u = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// The DST-I matrix S[a][b] = 2 sin(pi (a+1)(b+1) / (N+1)) of length N as two
// bf16 planes (hi = bf16(S), lo = bf16(S - hi)), Np x Np, zero padded.
static int get_sines(Ctx* c, int N, const __bf16** Sh, const __bf16** Sl) {
    auto it = c->sines.find(N);
    if (it == c->sines.end()) {
        const int Np = gemm_np_host(N), n = N + 1;
        std::vector<__bf16> h((size_t)2 * Np * Np, host_bf16(0.0f));
        const double pi = 4. * std::atan(1.0);
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b) {
                const long m = (long)(a + 1) * (b + 1) % (2 * n);   // exact argument reduction
                const double v = 2. * std::sin(pi * (double)m / n);
                const __bf16 hi = host_bf16((float)v);
                h[(size_t)a * Np + b] = hi;
                h[(size_t)Np * Np + (size_t)a * Np + b] = host_bf16((float)(v - (double)host_float(hi)));
            }
        __bf16* d = nullptr;
        CKR(c->mem.alloc(&d, h.size()));
        CK(hipMemcpy(d, h.data(), sizeof(__bf16) * h.size(), hipMemcpyHostToDevice));
        it = c->sines.emplace(N, std::make_pair(d, d + (size_t)Np * Np)).first;
    }
    *Sh = it->second.first;
    *Sl = it->second.second;
    return 0;
}

// Shape of the plan of a DST-I of length N: FFT length M = 2(N+1), its
// Stockham radices and the row pairs per workgroup (table pointers unset).
static int plan_shape(int N, FftPlan* P) {
    *P = FftPlan{};
    P->N = N;
    P->M = 2 * (N + 1);
    if (P->M > kFftLds) { set_error("DST length %d exceeds the LDS plan limit (%d)", N, kFftLds / 2 - 1); return -2; }
    int m = P->M;
    while (m % 8 == 0) { P->rad[P->nst++] = 8; m /= 8; }
    while (m % 4 == 0) { P->rad[P->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { P->rad[P->nst++] = 2; m /= 2; }
    while (m % 5 == 0) { P->rad[P->nst++] = 5; m /= 5; }
    while (m % 3 == 0) { P->rad[P->nst++] = 3; m /= 3; }
    if (m != 1) { set_error("DST length %d: 2(N+1) has a prime factor > 5", N); return -2; }
    P->np = std::max(1, std::min(kFftNpMax, kFftLds / P->M));
    return 0;
}

// Whether a plan of FFT length M has the second-stage table tw16.
static bool plan_has_tw16(int M) { return M == 1024 || M == 2048; }

// Host values of a plan's twiddle tables: tw (M entries) and, where
// plan_has_tw16, t16 (256 entries).
static void plan_tables(int M, double2* tw, double2* t16) {
    const double twopi = 8.0 * std::atan(1.0);
    for (int k = 0; k < M; ++k) {
        const double a = -twopi * (double)k / (double)M;
        tw[k] = make_double2(std::cos(a), std::sin(a));
    }
    if (plan_has_tw16(M)) {   // second-stage table of the register-held transforms
        const int tstep = M / 256;
        for (int r = 0; r < 16; ++r)
            for (int jm = 0; jm < 16; ++jm) t16[16 * r + jm] = tw[(tstep * jm * r) % M];
    }
}

static int get_plan(Ctx* c, int N, FftPlan* out) {
    auto it = c->plans.find(N);
    if (it != c->plans.end()) { *out = it->second; return 0; }
    FftPlan P{};
    CKR(plan_shape(N, &P));
    std::vector<double2> tw(P.M), t16(256);
    plan_tables(P.M, tw.data(), t16.data());
    double2* d = nullptr;
    CKR(c->mem.alloc(&d, P.M));
    CK(hipMemcpy(d, tw.data(), sizeof(double2) * P.M, hipMemcpyHostToDevice));
    P.tw = d;
    P.tw16 = nullptr;
    if (plan_has_tw16(P.M)) {
        double2* d16 = nullptr;
        CKR(c->mem.alloc(&d16, 256));
        CK(hipMemcpy(d16, t16.data(), sizeof(double2) * 256, hipMemcpyHostToDevice));
        P.tw16 = d16;
    }
    c->plans[N] = P;
    *out = P;
    return 0;
}

// ------------------------------------------------------------ solve plans
// One multi-domain solve of one flow.
struct SolveSpec {
    int nx, ny, L;
    double dx;             // finest spacing
    int kind;
    double alpha;
    const double* f;
    double* u;
    double* R;             // coarsify buffer (L levels)
    double* T1;
    double* T2;
    double in_scale;
    int epilogue;
    const double* W;
    const double* cosx;
    const double* cosy;
    const double* bcring;  // single-level explicit boundary values (L == 1 only)
    const double* skip;    // null, or the flow's force-solve flags
    double* vside;         // epilogue solves: side buffer (L levels) for the BC values of v
    double* u2 = nullptr;  // null, or a second output (L levels) of the same values
};

// Optional per-launch timing: a HIP event pair around every kernel launch on
// the batch stream, accumulated per kernel class with its algorithmic bytes.
// fs_*: the DST passes and coarsify of the M applications inside the force
// solve (kernels k_fs_*), counted with the bytes of the flows still iterating.
// ell32_* / fs32_*: the same passes with FP32 transforms (precision modes).
enum KClass { K_PASS1, K_PASS2, K_PASS3, K_COARSIFY, K_VEL, K_CROSS, K_RHS, K_CURLQ, K_INTERP,
              K_SMEAR, K_MINV, K_CG, K_FORCES, K_CLEAR, K_FSPASS1, K_FSPASS2, K_FSPASS3, K_FSCOARSIFY,
              K_FSHELM, K_HELM, K_PASS32_1, K_PASS32_2, K_PASS32_3, K_FSPASS32_1, K_FSPASS32_2,
              K_FSPASS32_3, K_GEMM, K_PCASM, K_PCNS, K_NCLASS };
static const char* kclass_names[K_NCLASS] = {"ell_pass1", "ell_pass2", "ell_pass3", "coarsify", "vel",
                                             "cross", "rhs", "curl_q", "interp", "smear_curl", "minv",
                                             "cg", "forces", "clear_nodes", "fs_pass1", "fs_pass2",
                                             "fs_pass3", "fs_coarsify", "fs_helm", "helm", "ell32_pass1",
                                             "ell32_pass2", "ell32_pass3", "fs32_pass1", "fs32_pass2",
                                             "fs32_pass3", "ell_gemm", "pc_assemble", "pc_refresh"};
struct Probe {
    struct Rec { int cls; hipEvent_t a, b; double bytes; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t ev() {
        if (used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    ~Probe() { for (auto e : pool) (void)hipEventDestroy(e); }
};

template <class Fn>
static void timed(Probe* pr, int cls, double bytes, hipStream_t s, Fn fn) {
    if (!pr) { fn(); return; }
    Probe::Rec r{cls, pr->ev(), pr->ev(), bytes};
    (void)hipEventRecord(r.a, s);
    fn();
    (void)hipEventRecord(r.b, s);
    pr->recs.push_back(r);
}

// The jobs of one level step that share their transform lengths (one launch
// per pass).
// Passes 1 and 3 over the whole level (EllJob's partial-pass fields off).
static void whole_level(EllJob& J) {
    J.full_lo = 0; J.full_hi = INT_MAX; J.fq0 = J.fq1 = -1;
    J.ring_lo = 1; J.ring_hi = 0; J.rj0 = J.rj1 = 1; J.ringsin = nullptr;
    J.sup_lo = 1; J.sup_hi = INT_MAX; J.sin1 = nullptr;
    J.rb_i0 = 0; J.rb_i1 = INT_MAX; J.rb_j0 = 0; J.rb_j1 = INT_MAX;
    J.t2_lo = 0; J.t2_hi = INT_MAX;
}

// The job of level lev of one solve (every field but the two FFT plans).
static void fill_ell_job(const SolveSpec& s, int lev, EllJob& J) {
    const long ns = (long)(s.nx - 1) * (s.ny - 1);
    const double dxl = s.dx * (1 << lev);
    whole_level(J);
    J.nx = s.nx; J.ny = s.ny; J.lev = lev; J.ngrid = s.L;
    J.kind = s.kind; J.epilogue = s.epilogue; J.in_scale = s.in_scale;
    J.alpha = s.alpha;
    J.bydx2 = 1. / (dxl * dxl);                                  // EllipticSolver2d.cc:44
    J.bc_coef = (s.kind == 0) ? 1 / (dxl * dxl) : s.alpha / (dxl * dxl);   // :137, :178
    J.norm = 1. / (2 * s.nx * 2 * s.ny);                         // :80
    J.f = s.f + lev * ns;
    J.R = s.R ? s.R + lev * ns : nullptr;
    // getBC source: the coarser level of the solution; for u = W - v
    // that is v, kept on the BC ring by the coarser level's pass 3
    const double* bcsrc = s.vside ? s.vside : s.u;
    J.uc = (lev < s.L - 1) ? bcsrc + (lev + 1) * ns : nullptr;
    J.vside = (s.vside && lev >= 1) ? s.vside + lev * ns : nullptr;
    J.bcring = (s.L == 1) ? s.bcring : nullptr;
    J.T1 = s.T1; J.T2 = s.T2;
    J.ld1 = t1_ld(s.nx - 1); J.ld2 = t1_ld(s.ny - 1);   // the scratch is sized for the padding
    J.u = s.u + lev * ns;
    J.u2 = s.u2 ? s.u2 + lev * ns : nullptr;
    J.W = s.W ? s.W + lev * ns : nullptr;
    J.cosx = s.cosx; J.cosy = s.cosy;
    J.skip = s.skip;
}

// Launch shape of the three passes of a single-job group: registers or LDS,
// workgroups and dynamic LDS bytes per pass (as build_solve_plan sizes them).
static void pass_shape(const EllJob& J, bool regs[3], int maxb[3], int lds[3]) {
    for (int k = 0; k < 3; ++k) {
        const long pairs = (k == 1 ? J.ny : J.nx) / 2;
        regs[k] = pass_in_registers(k == 1 ? J.pi.M : J.pj.M, pairs);
        maxb[k] = blocks_for_pass(J, k + 1, regs[k]);
        lds[k] = lds_for_pass(J, k + 1, regs[k]);
    }
}

struct PassGroup {
    EllJob* d = nullptr;
    int n = 0;
    int mj = 0, mi = 0;                             // FFT lengths along j (passes 1, 3) and i (pass 2)
    int maxb[3] = {0, 0, 0}, lds[3] = {0, 0, 0};
    bool regs[3] = {false, false, false};           // pass_in_registers per pass
    bool fp32 = false;                              // FP32 transforms (LDS passes, k32_*)
    bool gemm = false;                              // split-BF16 GEMM-DST on MFMA (launch_gemm_dst)
    const __bf16 *Sh = nullptr, *Sl = nullptr;      // its sine matrix
    int gemm_n = 0;                                 // its N (square levels only)
    double t_radix_us = 0, t_gemm_us = 0;           // autotuner timings of the group (0: not timed)
    int maxb64[3] = {0, 0, 0}, lds64[3] = {0, 0, 0};   // the FP64 launch shape, kept for switching back
    bool regs64[3] = {false, false, false};
    std::vector<int> jflow;                         // flow of each job
    std::vector<double> jbytes[3];                  // per pass: 16 B per point (one read + one write),
    double bytes[3] = {0, 0, 0};                    // less where a pass is restricted (pass_bytes)
    std::vector<EllJob> h;                          // host copy of the jobs
};

struct SolvePlan {
    bool force_solve = false;                       // M application (k_fs_* kernels)
    bool fp32 = false;                              // every group's transforms in FP32 (set_plan_fp32)
    int nlev = 0;                                   // max L over flows
    std::vector<std::vector<int>> cflow;            // flow of each coarsify job, per level
    std::vector<std::vector<double>> cjbytes;
    std::vector<double> cbytes;                     // per coarsify level
    std::vector<CoarsifyJob*> cj;                   // lev = 1..nlev-1
    std::vector<int> cj_n, cj_rows;   // coarsify jobs and the most coarse rows of a job, per level
    std::vector<std::vector<PassGroup>> steps;      // step s = 0..nlev-1 (coarsest level first)
};

static int build_solve_plan(Ctx* c, DevMem& mem, const std::vector<SolveSpec>& specs, SolvePlan& P) {
    P.nlev = 0;
    for (auto& s : specs) P.nlev = std::max(P.nlev, s.L);
    P.cj.assign(P.nlev, nullptr);
    P.cj_n.assign(P.nlev, 0);
    P.cj_rows.assign(P.nlev, 0);
    P.cflow.assign(P.nlev, {});
    P.cjbytes.assign(P.nlev, {});
    for (int lev = 1; lev < P.nlev; ++lev) {
        std::vector<CoarsifyJob> jobs;
        int pts = 0;
        for (size_t k = 0; k < specs.size(); ++k) {
            const SolveSpec& s = specs[k];
            if (lev >= s.L) continue;
            jobs.push_back(CoarsifyJob{s.nx, s.ny, lev, s.f, s.R, s.skip});
            pts = std::max(pts, s.nx / 2 - 1);
            P.cflow[lev].push_back((int)k);
            P.cjbytes[lev].push_back(8.0 * 5 * (s.nx / 2 - 1) * (s.ny / 2 - 1));
        }
        CoarsifyJob* d = nullptr;
        CKR(mem.alloc(&d, jobs.size()));
        CK(hipMemcpy(d, jobs.data(), sizeof(CoarsifyJob) * jobs.size(), hipMemcpyHostToDevice));
        P.cj[lev] = d;
        P.cj_n[lev] = (int)jobs.size();
        P.cj_rows[lev] = pts;
    }
    P.steps.assign(P.nlev, {});
    P.cbytes.assign(P.nlev, 0.0);
    for (int lev = 1; lev < P.nlev; ++lev)
        for (auto& s : specs)
            if (lev < s.L) P.cbytes[lev] += 8.0 * 5 * (s.nx / 2 - 1) * (s.ny / 2 - 1);   // 4 fine reads + 1 write
    for (int st = 0; st < P.nlev; ++st) {
        std::vector<PassGroup>& groups = P.steps[st];
        std::vector<std::vector<EllJob>> gjobs;
        for (size_t k = 0; k < specs.size(); ++k) {
            const SolveSpec& s = specs[k];
            if (st >= s.L) continue;
            const int lev = s.L - 1 - st;
            const long ns = (long)(s.nx - 1) * (s.ny - 1);
            EllJob J{};
            fill_ell_job(s, lev, J);
            CKR(get_plan(c, s.ny - 1, &J.pj));
            CKR(get_plan(c, s.nx - 1, &J.pi));
            size_t g = 0;
            while (g < groups.size() && !(groups[g].mj == J.pj.M && groups[g].mi == J.pi.M)) ++g;
            if (g == groups.size()) {
                groups.emplace_back();
                groups.back().mj = J.pj.M;
                groups.back().mi = J.pi.M;
                gjobs.emplace_back();
            }
            for (int q = 0; q < 3; ++q) groups[g].bytes[q] += 16.0 * ns;
            groups[g].jflow.push_back((int)k);
            for (int q = 0; q < 3; ++q) groups[g].jbytes[q].push_back(16.0 * ns);
            gjobs[g].push_back(J);
        }
        for (size_t g = 0; g < groups.size(); ++g) {
            PassGroup& G = groups[g];
            std::vector<EllJob>& jobs = gjobs[g];
            for (int k = 0; k < 3; ++k) {
                long pairs = 0;
                for (auto& J : jobs) pairs += (k == 1 ? J.ny : J.nx) / 2;   // ceil(rows / 2)
                G.regs[k] = pass_in_registers(k == 1 ? G.mi : G.mj, pairs);
                for (auto& J : jobs) {
                    G.maxb[k] = std::max(G.maxb[k], blocks_for_pass(J, k + 1, G.regs[k]));
                    G.lds[k] = std::max(G.lds[k], lds_for_pass(J, k + 1, G.regs[k]));
                }
            }
            for (int k = 0; k < 3; ++k) {
                G.regs64[k] = G.regs[k];
                G.maxb64[k] = G.maxb[k];
                G.lds64[k] = G.lds[k];
            }
            CKR(mem.alloc(&G.d, jobs.size()));
            CK(hipMemcpy(G.d, jobs.data(), sizeof(EllJob) * jobs.size(), hipMemcpyHostToDevice));
            G.n = (int)jobs.size();
            G.h = jobs;
        }
    }
    return 0;
}

static void group_bytes(PassGroup& G);

// Switches a plan's transforms between FP64 (the register-held or LDS
// passes chosen at build) and FP32 (LDS passes over every row: the partial-
// pass restrictions of register-held passes do not apply, so an FP32 pass
// computes a superset of what the FP64 one writes).
static void set_plan_fp32(SolvePlan& P, bool on) {
    P.fp32 = on;
    for (auto& step : P.steps)
        for (auto& G : step) {
            G.fp32 = on;
            G.gemm = false;
            for (int k = 0; k < 3; ++k) {
                if (!on) {
                    G.regs[k] = G.regs64[k];
                    G.maxb[k] = G.maxb64[k];
                    G.lds[k] = G.lds64[k];
                    continue;
                }
                G.regs[k] = false;
                G.maxb[k] = 0;
                G.lds[k] = 0;
                for (auto& J : G.h) {
                    G.maxb[k] = std::max(G.maxb[k], blocks_for_pass(J, k + 1, false));
                    G.lds[k] = std::max(G.lds[k], lds_for_pass32(J, k + 1));
                }
            }
            if (!G.jbytes[0].empty()) group_bytes(G);
        }
}

// Whether a group can take the GEMM-DST: square levels of N <= 511 whose
// transposed scratch holds the two bf16 plane pairs (2 Np^2 each).
static bool gemm_eligible(const PassGroup& G) {
    if (G.h.empty()) return false;
    for (const EllJob& J : G.h) {
        const int N = J.nx - 1, Np = gemm_np_host(N);
        if (J.nx != J.ny || N > 511 || N != G.h[0].nx - 1) return false;
        if (4.0 * Np * Np > 8.0 * (double)(J.ny - 1) * t1_ld(J.nx - 1)) return false;
    }
    return true;
}

// The reduced-precision plan's transform per group: mode 3 times the FP32
// radix passes and the split-BF16 GEMM-DST on the group's own jobs (outputs
// sent to scratch, so no field changes) and keeps the faster; mode 4 takes the
// GEMM wherever it is eligible.  Groups that are not eligible keep the FP32
// radix passes.  The plan must already be FP32 (set_plan_fp32).
static int tune_plan_gemm(Ctx* c, DevMem& mem, SolvePlan& P, int mode) {
    hipStream_t s = c->stream;
    for (auto& step : P.steps)
        for (auto& G : step) {
            G.gemm = false;
            if (!gemm_eligible(G)) continue;
            const int N = G.h[0].nx - 1;
            CKR(get_sines(c, N, &G.Sh, &G.Sl));
            G.gemm_n = N;
            if (mode == 4) { G.gemm = true; continue; }
            std::vector<EllJob> jobs = G.h;
            long maxns = 0;
            for (auto& J : jobs) maxns = std::max(maxns, (long)(J.nx - 1) * (J.ny - 1));
            double* out = nullptr;
            EllJob* d = nullptr;
            CKR(mem.alloc(&out, maxns));
            for (auto& J : jobs) {   // write to scratch; no epilogue, side outputs or skip flags
                J.u = out; J.u2 = nullptr; J.vside = nullptr; J.epilogue = 0; J.W = nullptr; J.skip = nullptr;
            }
            CKR(mem.alloc(&d, jobs.size()));
            CK(hipMemcpy(d, jobs.data(), sizeof(EllJob) * jobs.size(), hipMemcpyHostToDevice));
            auto radix = [&] {
                for (int pass = 1; pass <= 3; ++pass)
                    launch_ell_pass32(pass, d, G.n, G.maxb[pass - 1], G.lds[pass - 1], s, P.force_solve,
                                      pass == 2 ? G.mi : G.mj);
            };
            auto gemm = [&] { launch_gemm_dst(d, G.n, N, G.Sh, G.Sl, s); };
            hipEvent_t e0, e1;
            CK(hipEventCreate(&e0));
            CK(hipEventCreate(&e1));
            float ms_r = 0, ms_g = 0;
            for (int rep = 0; rep < 2; ++rep) {   // a warm-up round, then the timed one
                CK(hipEventRecord(e0, s));
                for (int k = 0; k < 3; ++k) radix();
                CK(hipEventRecord(e1, s));
                CK(hipEventSynchronize(e1));
                CK(hipEventElapsedTime(&ms_r, e0, e1));
                CK(hipEventRecord(e0, s));
                for (int k = 0; k < 3; ++k) gemm();
                CK(hipEventRecord(e1, s));
                CK(hipEventSynchronize(e1));
                CK(hipEventElapsedTime(&ms_g, e0, e1));
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            CK(hipGetLastError());
            G.t_radix_us = ms_r * 1e3 / 3;
            G.t_gemm_us = ms_g * 1e3 / 3;
            G.gemm = G.t_gemm_us < G.t_radix_us;
        }
    return 0;
}

// The fine-to-coarse chain of a solve's right-hand side (central regions of
// levels >= 1 into R).  active (probe only): flows whose jobs do work; the
// others are skipped by their flags, so their bytes are not counted.
static void run_coarsify(const SolvePlan& P, hipStream_t s, Probe* pr, const std::vector<char>* active) {
    auto sum = [&](const std::vector<int>& fl, const std::vector<double>& by, double all) {
        if (!active) return all;
        double b = 0;
        for (size_t j = 0; j < fl.size(); ++j)
            if ((*active)[fl[j]]) b += by[j];
        return b;
    };
    const bool fs = P.force_solve;
    for (int lev = 1; lev < P.nlev; ++lev)
        timed(pr, fs ? K_FSCOARSIFY : K_COARSIFY, pr ? sum(P.cflow[lev], P.cjbytes[lev], P.cbytes[lev]) : 0, s,
              [&] { launch_coarsify(P.cj[lev], P.cj_n[lev], P.cj_rows[lev], s, fs); });
}

// active (probe only): flows whose jobs do work; the others are skipped by
// their flags, so their bytes are not counted.
// coarsify = false: the caller has built the coarse right-hand sides.
static void run_solve(const SolvePlan& P, hipStream_t s, Probe* pr = nullptr,
                      const std::vector<char>* active = nullptr, bool coarsify = true) {
    auto sum = [&](const std::vector<int>& fl, const std::vector<double>& by, double all) {
        if (!active) return all;
        double b = 0;
        for (size_t j = 0; j < fl.size(); ++j)
            if ((*active)[fl[j]]) b += by[j];
        return b;
    };
    const bool fs = P.force_solve;
    if (coarsify) run_coarsify(P, s, pr, active);
    for (int st = 0; st < P.nlev; ++st)
        for (const PassGroup& G : P.steps[st]) {
            if (G.gemm) {
                double by = 0;
                if (pr)
                    for (size_t j = 0; j < G.jflow.size(); ++j)
                        if (!active || (*active)[G.jflow[j]]) by += 48.0 * (G.h[j].nx - 1) * (G.h[j].ny - 1);
                timed(pr, K_GEMM, by, s, [&] { launch_gemm_dst(G.d, G.n, G.gemm_n, G.Sh, G.Sl, s); });
                continue;
            }
            for (int pass = 1; pass <= 3; ++pass) {
                const int cls = G.fp32 ? (fs ? K_FSPASS32_1 : K_PASS32_1) : (fs ? K_FSPASS1 : K_PASS1);
                timed(pr, cls + pass - 1, pr ? sum(G.jflow, G.jbytes[pass - 1], G.bytes[pass - 1]) : 0, s, [&] {
                    if (G.fp32)
                        launch_ell_pass32(pass, G.d, G.n, G.maxb[pass - 1], G.lds[pass - 1], s, fs,
                                          pass == 2 ? G.mi : G.mj);
                    else
                        launch_ell_pass(pass, G.d, G.n, G.maxb[pass - 1], G.lds[pass - 1], s, fs,
                                        pass == 2 ? G.mi : G.mj, G.regs[pass - 1]);
                });
            }
        }
}

// cos(pi i / n), i = 1..n-1 (EllipticSolver2d.cc:47-48) into t[0..n-2].
static void cos_fill(int n, double* t) {
    const double pi = 4. * std::atan(1.0);
    for (int i = 1; i < n; ++i) t[i - 1] = std::cos((pi * i) / n);
}

static int cos_tables(DevMem& mem, int n, double** out) {
    std::vector<double> t(n - 1);
    cos_fill(n, t.data());
    CKR(mem.alloc(out, n - 1));
    CK(hipMemcpy(*out, t.data(), sizeof(double) * (n - 1), hipMemcpyHostToDevice));
    return 0;
}

// ---------------------------------------------------------------- batch
struct Flow {
    ibpm_flow_params p{};
    std::string geom_text;
    Geometry geom;
    int nx, ny, L, nb;
    double dx;
    long ns, nf;
    bool cg = false;
    // device fields
    double *omega, *q, *nprev, *a, *wstar, *psi, *c, *U, *V, *cross, *R, *T1, *T2, *mt, *cosx, *cosy;
    double *f, *rhs, *b, *cg_r, *cg_d, *cg_q, *cg_state, *hist, *unit, *mcol;
    // models: base states (linear models; one per phase of the periodic one),
    // the adjoint's Flux x Flux cross product, SFD's filtered state
    std::vector<double*> bq, bw;
    std::vector<bool> base_set;
    double *sff = nullptr, *what = nullptr, *sfd_prev = nullptr;
    double base_mag = 1.0;        // magnitude of the uniform base flow (0: no potential flow)
    int* hist_count;
    double* minv[4] = {nullptr, nullptr, nullptr, nullptr};
    // preconditioned CG: dense P, its input/output, the Krylov pairs of the
    // last solve of each substep (double-buffered) and their counts
    double *pinv = nullptr, *cg_t = nullptr, *cg_u = nullptr, *lmp_g = nullptr, *lmp_c = nullptr;
    double *lmp_s[4][2] = {}, *lmp_y[4][2] = {};
    int* lmp_cnt = nullptr;
    double pbeta = 0.0;            // beta of the substep P was built for
    // P from the kernel table (pc_build_levels 2, 3): G's responses per
    // substep (shared by the flows of one grid, Re and dt), and the product
    // buffer of P's Newton-Schulz refresh
    const double* ktab[4] = {nullptr, nullptr, nullptr, nullptr};
    const int* ksrc = nullptr;     // the table's unit-flux nodes (i, j), nsrc of them
    int nsrc = 0;
    double bbox0[4] = {0, 0, 0, 0};   // the points' extent at set-up: x min, max, y min, max
    double *pc_T = nullptr, *pc_X = nullptr;
    std::string chol_cache;        // basename of the .cholesky factor files ("" = none)
    int chol_loaded = 0;           // substep factors read from those files
    // regulariser
    RegLists reg;
    int *e_ptr, *e_f, *t_ptr, *t_k, *node_idx, *old_idx;
    double *e_w, *t_w;
    double* e_sep;                 // E rows in separable form (kESep per component)
    int4* node_nbr;
    int nnodes = 0, nold = 0;
    // algorithmic bytes of one k_interp (without the constraint b) and one
    // k_smear_curl launch over this flow's current lists (reg_bytes)
    double by_interp = 0, by_smear = 0;
    std::vector<int> reg_old;      // node list of the previous association (cleared on the device)
    int reg_old_n = 0;
    // the next substep's kinematics, regulariser lists and constraint, computed
    // while the GPU runs this substep (kinematics_ahead), with the motions'
    // running state from before that move (restored if they go unused)
    struct {
        bool valid = false;
        double t = 0;
        std::vector<double> px, py, pu, pv, b;
        RegLists reg;
        std::vector<std::array<double, 4>> state;
    } ahead;
    std::vector<double> b_host;    // the constraint taken from `ahead` for upload_constraint
    bool b_host_ready = false;
    HostStage stage;
    // local Helmholtz path of the force solve's A^-1: result and sweep buffer
    // (all levels) and, per level, the box they were last swept over (zero off it)
    double *m1 = nullptr, *m1tmp = nullptr;
    double* ringsin = nullptr;     // 2 sin(pi l j / ny) at the ring columns (restrict_m2l_pass3)
    double* sin1 = nullptr;        // 2 sin(pi k / ny), k = 1..ny-1
    std::vector<std::array<int, 4>> hbox;
    // host state
    std::vector<double> px, py, pu, pv;
    int hist_cap = 0;
    int cg_iters[4] = {-1, -1, -1, -1};
    // learned warm start: ring of the final force and the constraint of the
    // last kRing substeps, features and scale of the current substep
    double *ring_f = nullptr, *ring_b = nullptr, *nn_x = nullptr, *nn_scale = nullptr;
    // unsteady base flow (ubf): the transferred motion, its u, v, w on the
    // device, and the angle that rotates the force after each step
    bool ubf = false;
    BaseFlowMotion bf;
    double* bf_dev = nullptr;
    double force_alpha = 0.0;
    std::vector<double> alpha_hist;
};

}  // namespace ibpmg

struct ibpm_ctx {
    ibpmg::Ctx c;
};

struct ibpm_batch {
    ibpm_ctx* ctx = nullptr;
    std::vector<ibpmg::Flow> flows;
    ibpmg::DevMem mem;
    int nsub = 3;
    double an[4], bn[4], cn[4];
    bool all_stationary = true, any_cg = false, initialized = false, any_ubf = false;
    bool any_periodic = false, any_adjoint = false;
    ibpmg::CrossFFOps* d_cff = nullptr;                // adjoint flows' Flux x Flux cross products
    int cff_n = 0;
    std::vector<ibpmg::CoarsifyJob*> d_cffc;          // their coarsify, per level
    std::vector<int> cffc_n;
    int cffc_rows = 0;                                 // coarse rows of the largest of them
    bool old_saved = false;
    double time = 0.0;
    int timestep = 0;
    long max_ns_all = 0, max_nf = 0, max_cells_all = 0;   // cells: (nx + 1)(ny + 1) L
    int max_comp = 0, max_nodes = 0, max_n = 0;
    // plans per substep
    ibpmg::SolvePlan S1[4], S2[4], S3[4], S4, M1[4], M2;
    // local Helmholtz path (plan_local_helm): on per substep, its jobs, sweeps
    // and the DST solve of L^-1 that reads its result
    bool helm_allowed = true, helm_on[4] = {false, false, false, false}, m2l_built = false;
    ibpmg::SolvePlan M2L;
    ibpmg::HelmJob* d_helm = nullptr;
    ibpmg::HelmJob* d_helm3 = nullptr;   // the same solves for omega = omega* - A^-1 c (no skip flags)
    // level-0 jobs of d_helm alone (the finest-level build of the preconditioner)
    ibpmg::HelmJob* d_helm0 = nullptr;
    int helm0_n = 0, helm0_blocks = 1;
    double helm0_bytes = 0;
    ibpmg::HelmJob* d_hclear = nullptr;  // boxes to clear when a flow's boxes change
    int hclear_cap = 0;
    int helm_cap = 0, helm_n = 0, helm_blocks = 0, helm_cblocks = 0, helm_ablocks = 0, helm_sweeps = 0;
    std::vector<int> helm_flow;
    std::vector<double> helm_bytes, helm_cbytes;
    // the L^-1 solve's coarsify chain over its right-hand side's support
    // (restrict_m2l_pass3), when every pass 1 of that solve is register-held
    bool m2l_boxed = false;
    int precision = 0;             // ibpm_batch_set_precision
    bool fs32 = false;             // the force solve's solves in FP32
    ibpmg::HelmJob* d_m2lc = nullptr;
    int m2lc_cap = 0, m2lc_n = 0, m2lc_blocks = 0;
    std::vector<int> m2lc_flow;
    std::vector<double> m2lc_bytes;
    long st_local_mapps = 0;
    ibpmg::FlowOps* d_fops[4] = {};
    ibpmg::RegOps *d_interp_c = nullptr, *d_interp_m = nullptr, *d_smear_f[4] = {}, *d_smear_d[4] = {};
    ibpmg::RegOps* d_clear = nullptr;
    ibpmg::MinvOps* d_minv[4] = {};
    ibpmg::CgOps* d_cg = nullptr;
    ibpmg::ForceOps* d_force = nullptr;
    std::vector<ibpmg::FlowOps> h_fops[4];
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    ibpmg::Probe* probe = nullptr;
    double bytes_vel = 0, bytes_cross = 0, bytes_rhs = 0, bytes_curlq = 0;
    int cg_chunk = 4;
    // learned warm start (learn.hip)
    int warm_mode = 1;            // 0 zero, 1 previous force (reference), 2 learned
    bool collect = false;
    long ksub = 0;                // substeps done since init
    int ring = 0;                 // ring length (2 S + 2)
    double* nn_params = nullptr;
    double *nn_m = nullptr, *nn_v = nullptr, *nn_grad = nullptr;
    long nn_adam_t = 0;
    bool nn_ready = false;
    double *data_x = nullptr, *data_y = nullptr;
    long data_cap = 0, data_n = 0;
    ibpmg::NnOps* d_nn = nullptr;
    ibpmg::NnOps* d_nn_inf = nullptr;   // nn_infer_table
    // force-solve state of every flow, 8 doubles each: [rho or delta, tol^2,
    // iterations, converged, broken down, pairs in use, r.r, spare]
    double* cg_states = nullptr;
    // preconditioned CG
    int pc_mode = 0, pc_pairs = 0;
    bool pc_valid = false, pc_broken = false;
    int pc_par[4] = {0, 0, 0, 0};
    int pc_base[4] = {-1, -1, -1, -1};   // iterations of the first solve after a build
    int pc_last[4] = {0, 0, 0, 0};
    double pc_excess = 0.0;              // iterations above pc_base since the build
    // P from M (0), from its finest-level form (1: M with L^-1 on level 0
    // alone, zero boundary values; plan M1L, build_inverse), or from the
    // kernel table (2; 3: and refreshed every substep by one Newton-Schulz
    // step against the table's M at the substep's points)
    int pc_build_levels = 0;
    // kernel tables: per grid class, the level-0 flux responses of G (the
    // flux-to-flux operator inside M, without beta) to a unit X and a unit Y
    // flux at the centre, per substep
    struct KernTab {
        int nx, ny, ngrid;
        double dx, re, dt;
        double* tab[4];                   // per substep: nsrc x (X response, Y response)
        std::vector<int> src;             // unit-flux nodes (i, j)
        int* d_src;
    };
    std::vector<KernTab> ktabs;
    ibpmg::PcAsmOps* d_pcasm[4] = {};     // the refresh's assembly per substep (scale pbeta)
    ibpmg::PcAsmOps* d_pcasm_b = nullptr; // a build's assembly
    // Newton-Schulz steps alternate between P and pc_X: d_ns1[0], d_ns2[0]
    // take P to pc_X, d_ns1[1], d_ns2[1] take pc_X back to P; an odd count
    // ends with d_ns_copy (pc_X to P)
    ibpmg::GemmOps *d_ns1[2] = {}, *d_ns2[2] = {};
    ibpmg::CopyPiece* d_ns_copy = nullptr;
    long st_refreshes = 0;
    // the refresh runs on a stream of its own beside the substep's right-hand
    // side and step solves: it starts after the substep's uploads (pc_ev_in)
    // and the force solve waits for it (pc_ev_out)
    hipStream_t pc_stream = nullptr;
    hipEvent_t pc_ev_in = nullptr, pc_ev_out = nullptr;
    bool pc_pending = false;
    // Newton-Schulz steps per refresh: with the multi-source table (spectrum
    // of M~^-1 M within 1e-4 of 1) a second step takes P's own error from
    // about 1e-3 to 1e-6, and every tandem solve from 3 iterations to 2
    int pc_ns_steps = 2;
    double pc_build_cost = 1.0;          // a build's cost in M applications per unknown (rebuild rule)
    ibpmg::SolvePlan M1L;
    bool m1l_built = false;
    // wall time of the last build and of a PCG iteration (moving average):
    // the rebuild rule's costs
    double pc_t_build = 0.0, pc_t_iter = 0.0;
    // build_inverse: interpolation descriptors of every column, unit vectors
    // substep prep uploads (moving bodies): asynchronous on the stream from
    // pinned buffers, alternate substeps; uev[p]: after the last copy made from
    // half p (the next reuse of half p waits for it)
    // solver polls: states copied to pinned memory behind the queued work
    double* h_states = nullptr;   // 8 per flow (poll_begin)
    double* h_reset = nullptr;    // reset_states' source
    hipEvent_t poll_ev = nullptr;
    ibpmg::HostStage ustage[2];
    unsigned char* d_ustage[2] = {nullptr, nullptr};   // device images of the halves
    std::vector<ibpmg::CopyPiece> upieces;              // this prep's pieces (async_flush)
    // k_copy_pieces tables: cg_d = f per flow (start_residual); the history
    // ring's f and b copies per ring slot
    ibpmg::CopyPiece *d_start_pieces = nullptr, *d_ring_pieces = nullptr;
    // zero fills of a zero start, one launch each: f of every flow, then M x0
    ibpmg::CopyPiece* d_zero_pieces = nullptr;
    hipEvent_t uev[2] = {nullptr, nullptr};
    bool upending[2] = {false, false};
    int upar = 0;
    bool async_up = false;
    ibpmg::RegOps* d_interp_cols = nullptr;
    size_t interp_cols_cap = 0;
    ibpmg::UnitOps* d_units = nullptr;
    ibpmg::PcgOps* d_pcg[4][2] = {};
    ibpmg::MinvOps* d_pminv[4] = {};
    // counters: iterations, solves, M applications in solves, M applications
    // in preconditioner builds, builds, fallbacks to plain CG
    long st_iters = 0, st_solves = 0, st_mapps = 0, st_build_mapps = 0, st_builds = 0, st_fallbacks = 0;
    double st_host_kin_s = 0, st_host_poll_s = 0;   // host time: kinematics + regulariser, solver polls
    double st_build_s = 0, st_build_host_s = 0;     // preconditioner builds: wall time, host factorisation
};

namespace ibpmg {

static int upload(void* d, const void* h, size_t bytes) {
    if (bytes == 0) return 0;
    CK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return 0;
}

// Starts a round of staged uploads for a flow: waits for copies still in
// flight from its staging buffer, then reuses it from the start.
static int stage_begin(hipStream_t s, HostStage& st) {
    if (st.pending) CK(hipStreamSynchronize(s));
    st.pending = false;
    st.used = 0;
    return 0;
}

// Copies host data into the staging buffer and enqueues its upload.
static int stage_upload(hipStream_t s, HostStage& st, void* d, const void* h, size_t bytes) {
    if (bytes == 0) return 0;
    const size_t off = (st.used + 255) & ~(size_t)255;
    if (off + bytes > st.cap) { set_error("upload staging buffer too small"); return -3; }
    std::memcpy(st.p + off, h, bytes);
    st.used = off + bytes;
    CK(hipMemcpyAsync(d, st.p + off, bytes, hipMemcpyHostToDevice, s));
    st.pending = true;
    return 0;
}

// Sends this prep's pieces: their data and the piece list in one copy of the
// pinned half to its device image, then one k_copy_pieces launch that puts
// every piece in place (instead of one runtime copy per piece).
static int async_flush(ibpm_batch* B) {
    if (B->upieces.empty()) return 0;
    hipStream_t s = B->ctx->c.stream;
    const int p = B->upar;
    HostStage& st = B->ustage[p];
    const size_t off = (st.used + 255) & ~(size_t)255, bytes = sizeof(CopyPiece) * B->upieces.size();
    if (off + bytes > st.cap) { set_error("upload staging buffer too small for the piece list"); return -3; }
    std::memcpy(st.p + off, B->upieces.data(), bytes);
    st.used = off + bytes;
    CK(hipMemcpyAsync(B->d_ustage[p], st.p, st.used, hipMemcpyHostToDevice, s));
    launch_copy_pieces((const CopyPiece*)(B->d_ustage[p] + off), (int)B->upieces.size(), s);
    B->upieces.clear();
    B->upending[p] = true;
    return 0;
}

// One upload of a substep's prep: its data into this substep's pinned half
// and a piece for async_flush, ordered on the stream after the kernels of
// earlier substeps that read the old contents.  Past the half's capacity
// (the piece list keeps 64 KB), the pieces so far are sent and this one is
// copied synchronously.
static int async_piece(ibpm_batch* B, void* d, const void* h, size_t bytes) {
    if (bytes == 0) return 0;
    HostStage& st = B->ustage[B->upar];
    const size_t off = (st.used + 255) & ~(size_t)255;
    if (bytes % 4 != 0 || off + bytes + ((size_t)64 << 10) > st.cap) {
        CKR(async_flush(B));
        CK(hipStreamSynchronize(B->ctx->c.stream));
        return upload(d, h, bytes);
    }
    std::memcpy(st.p + off, h, bytes);
    st.used = off + bytes;
    B->upieces.push_back(CopyPiece{(const unsigned int*)(B->d_ustage[B->upar] + off), (unsigned int*)d,
                                   (long)(bytes / 4)});
    return 0;
}

// A descriptor upload: during a moving-body substep's prep (async_up) an
// async_piece, else a synchronous copy.
static int up(ibpm_batch* B, void* d, const void* h, size_t bytes) {
    return B->async_up ? async_piece(B, d, h, bytes) : upload(d, h, bytes);
}

// Starts a moving-body substep's prep: the other pinned half, once the copies
// made from it two substeps ago have run.
static int async_prep_begin(ibpm_batch* B) {
    if (!B->ustage[0].p) {
        for (int p = 0; p < 2; ++p) {
            B->ustage[p].cap = (size_t)16 << 20;
            CK(hipHostMalloc((void**)&B->ustage[p].p, B->ustage[p].cap, hipHostMallocDefault));
            CKR(B->mem.alloc(&B->d_ustage[p], B->ustage[p].cap));
        }
        for (auto& e : B->uev) CK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    B->upar ^= 1;
    const int p = B->upar;
    if (B->upending[p]) CK(hipEventSynchronize(B->uev[p]));
    B->upending[p] = false;
    B->ustage[p].used = 0;
    B->ustage[p].pending = false;
    B->upieces.clear();
    B->async_up = true;
    return 0;
}

static int async_prep_end(ibpm_batch* B) {
    CKR(async_flush(B));
    B->async_up = false;
    CK(hipEventRecord(B->uev[B->upar], B->ctx->c.stream));
    B->upending[B->upar] = true;
    return 0;
}

// Scheme coefficients, Scheme.h:14-56
static void scheme_coefs(int scheme, ibpm_batch* B) {
    switch (scheme) {
    case 0: B->nsub = 1; B->an[0] = 1.; B->bn[0] = 0.; B->cn[0] = 1.; break;
    case 1: B->nsub = 1; B->an[0] = 3. / 2.; B->bn[0] = -1. / 2.; B->cn[0] = 1.; break;
    case 3:
        B->nsub = 4;
        B->an[0] = 8. / 17.; B->an[1] = 17. / 60.; B->an[2] = 5. / 12.; B->an[3] = 3. / 4.;
        B->bn[0] = 0.; B->bn[1] = -15. / 68.; B->bn[2] = -17. / 60.; B->bn[3] = -5. / 12.;
        B->cn[0] = 8. / 17.; B->cn[1] = 8. / 15.; B->cn[2] = 2. / 3.; B->cn[3] = 1.;
        break;
    default:
        B->nsub = 3;
        B->an[0] = 8. / 15.; B->an[1] = 5. / 12.; B->an[2] = 3. / 4.;
        B->bn[0] = 0.; B->bn[1] = -17. / 60.; B->bn[2] = -5. / 12.;
        B->cn[0] = 8. / 15.; B->cn[1] = 2. / 3.; B->cn[2] = 1.;
    }
}

static double beta_of(const ibpm_batch* B, const Flow& F, int i) { return (B->an[i] + B->bn[i]) * F.p.dt; }

// The constraint b = u_body - E q0 (NavierStokesModel.cc:86-91,142-146) with
// q0 on level 0: the uniform flow, or the unsteady base flow at its last
// update (Flux.cc:101-129 at the edge coordinates of Grid.cc:104-139).
static void compute_constraint(const Flow& F, const RegLists& R, const std::vector<double>& pu,
                               const std::vector<double>& pv, std::vector<double>& b) {
    const double pi = 4. * std::atan(1.);
    const double al = F.p.alpha_deg * pi / 180.;
    double sal, cal;
    sin_cos(al, sal, cal);
    const double u0 = F.base_mag * cal * F.dx, v0 = F.base_mag * sal * F.dx;
    const long nfx = (long)(F.nx + 1) * F.ny;
    auto q0 = [&](int fidx) {
        if (!F.ubf) return fidx < nfx ? u0 : v0;
        if (fidx < nfx) {   // X flux (i, j): x edge i, y centre j
            const int i = fidx / F.ny, j = fidx % F.ny;
            return F.bf.flux(0, F.p.xoffset + i * F.dx, F.p.yoffset + (j + 0.5) * F.dx, F.dx);
        }
        const int u = fidx - (int)nfx, i = u / (F.ny + 1), j = u % (F.ny + 1);
        return F.bf.flux(1, F.p.xoffset + (i + 0.5) * F.dx, F.p.yoffset + j * F.dx, F.dx);
    };
    b.assign(2 * F.nb, 0.0);
    for (int k = 0; k < 2 * F.nb; ++k) {
        double s = 0;
        for (int a = R.e_ptr[k]; a < R.e_ptr[k + 1]; ++a) s += R.e_w[a] * q0(R.e_f[a]);
        s *= 1.0 / F.dx;
        const double vel = k < F.nb ? pu[k] : pv[k - F.nb];
        b[k] = vel - s;
    }
}

// Uploads the constraint of F's current lists and velocities (computed ahead
// by kinematics_ahead where it took that route).
static void take_constraint(Flow& F, std::vector<double>& b) {
    if (F.b_host_ready) {
        b.swap(F.b_host);
        F.b_host_ready = false;
    } else {
        compute_constraint(F, F.reg, F.pu, F.pv, b);
    }
}

static int upload_constraint(hipStream_t s, Flow& F, HostStage& st) {
    std::vector<double> b;
    take_constraint(F, b);
    return stage_upload(s, st, F.b, b.data(), sizeof(double) * b.size());
}

// BaseFlow::moveFlow(t) of a flow with an unsteady base flow
// (NavierStokesModel::updateOperators, NavierStokesModel.cc:93-99): the new
// u, v, w on the device and the constraint for them.
static int move_base_flow(ibpm_batch* B, Flow& F, double t) {
    hipStream_t s = B->ctx->c.stream;
    F.bf.move(t);
    CKR(stage_begin(s, F.stage));
    const double uvw[3] = {F.bf.u, F.bf.v, F.bf.w};
    CKR(stage_upload(s, F.stage, F.bf_dev, uvw, sizeof uvw));
    return upload_constraint(s, F, F.stage);
}

// Regulariser lists of a flow for its current points; uploads them and the
// constraint b = u_body - E q0.
// Host part: the lists for the flow's current points; the previous node list
// is kept so the device can clear those nodes.
static void build_regularizer(Flow& F) {
    F.reg_old = F.reg.node_idx;
    F.reg_old_n = F.nnodes;
    build_reg_lists(F.nx, F.ny, F.dx, F.p.xoffset, F.p.yoffset, F.px, F.py, F.reg);
}

// The running state of the motions that integrate over successive calls.
static void save_motion_state(const Flow& F, std::vector<std::array<double, 4>>& st) {
    st.clear();
    for (const Body& b : F.geom.bodies) st.push_back({b.acc[0], b.acc[1], b.acc[2], b.oldtime});
}

static void restore_motion_state(Flow& F, const std::vector<std::array<double, 4>>& st) {
    for (size_t k = 0; k < st.size() && k < F.geom.bodies.size(); ++k) {
        const Body& b = F.geom.bodies[k];
        b.acc[0] = st[k][0]; b.acc[1] = st[k][1]; b.acc[2] = st[k][2]; b.oldtime = st[k][3];
    }
}

// A moving flow's kinematics, regulariser lists and constraint at time t,
// ahead of the substep that uses them (the motion is moved now, in the
// reference's sequence of times; its prior state is kept).
static void kinematics_ahead(Flow& F, double t) {
    auto& A = F.ahead;
    save_motion_state(F, A.state);
    F.geom.move(t, A.px, A.py, A.pu, A.pv);
    build_reg_lists(F.nx, F.ny, F.dx, F.p.xoffset, F.p.yoffset, A.px, A.py, A.reg);
    compute_constraint(F, A.reg, A.pu, A.pv, A.b);
    A.t = t;
    A.valid = true;
}

// A moving flow's kinematics for its substep at time t: those computed ahead
// for t, or computed now (after undoing a move made ahead for another time).
static void kinematics_at(Flow& F, double t) {
    auto& A = F.ahead;
    if (A.valid && A.t == t) {
        F.reg_old = F.reg.node_idx;
        F.reg_old_n = F.nnodes;
        std::swap(F.reg, A.reg);
        F.px.swap(A.px); F.py.swap(A.py); F.pu.swap(A.pu); F.pv.swap(A.pv);
        F.b_host.swap(A.b);
        F.b_host_ready = true;
        A.valid = false;
        return;
    }
    if (A.valid) {
        restore_motion_state(F, A.state);
        A.valid = false;
    }
    F.geom.move(t, F.px, F.py, F.pu, F.pv);
    build_regularizer(F);
    F.b_host_ready = false;
}

// Undoes every move made ahead (the batch's time or bodies are set anew).
static void cancel_ahead(ibpm_batch* B) {
    for (auto& F : B->flows)
        if (F.ahead.valid) {
            restore_motion_state(F, F.ahead.state);
            F.ahead.valid = false;
        }
}

static int upload_regularizer(ibpm_batch* B, Flow& F) {
    const RegLists& R = F.reg;
    const std::vector<int>& old = F.reg_old;
    const int old_nodes = F.reg_old_n;
    const int cap = 18 * F.nb + 1;
    if ((int)R.e_f.size() > cap || (int)R.t_k.size() > cap || (int)R.node_idx.size() > 4 * cap) {
        set_error("regulariser list exceeds capacity");
        return -3;
    }
    hipStream_t s = B->ctx->c.stream;
    // during a substep's prep: pieces of its pinned half (async_piece); else
    // the flow's own buffer, once its last copies have run
    HostStage& st = F.stage;
    if (!B->async_up) CKR(stage_begin(s, st));
    auto put = [&](void* d, const void* h, size_t bytes) {
        return B->async_up ? async_piece(B, d, h, bytes) : stage_upload(s, st, d, h, bytes);
    };
    CKR(put(F.e_ptr, R.e_ptr.data(), sizeof(int) * R.e_ptr.size()));
    CKR(put(F.e_f, R.e_f.data(), sizeof(int) * R.e_f.size()));
    CKR(put(F.e_w, R.e_w.data(), sizeof(double) * R.e_w.size()));
    CKR(put(F.e_sep, R.e_sep.data(), sizeof(double) * R.e_sep.size()));
    CKR(put(F.t_ptr, R.t_ptr.data(), sizeof(int) * R.t_ptr.size()));
    CKR(put(F.t_k, R.t_k.data(), sizeof(int) * R.t_k.size()));
    CKR(put(F.t_w, R.t_w.data(), sizeof(double) * R.t_w.size()));
    CKR(put(F.node_idx, R.node_idx.data(), sizeof(int) * R.node_idx.size()));
    CKR(put(F.node_nbr, R.node_nbr.data(), sizeof(int) * R.node_nbr.size()));
    CKR(put(F.old_idx, old.data(), sizeof(int) * old.size()));
    F.nold = old_nodes;
    F.nnodes = (int)R.node_idx.size();
    // k_interp: per E entry its flux index, weight and two psi values (28 B),
    // per component the row pointer and the output (12 B); k_smear_curl: per
    // node its four slots, scalar index and output (28 B), per slot used the
    // row pointers (8 B) and per E^T entry its component, weight and f (20 B)
    F.by_interp = 28.0 * (double)R.e_f.size() + 12.0 * (double)(2 * F.nb);
    double sm = 28.0 * (double)F.nnodes;
    for (int slot : R.node_nbr)
        if (slot >= 0) sm += 8.0 + 20.0 * (double)(R.t_ptr[slot + 1] - R.t_ptr[slot]);
    F.by_smear = sm;
    std::vector<double> b;
    take_constraint(F, b);
    return put(F.b, b.data(), sizeof(double) * b.size());
}

static int update_regularizer(ibpm_batch* B, Flow& F) {
    build_regularizer(F);
    return upload_regularizer(B, F);
}

// Runs body(q) for q < n on up to 16 host threads (q = tid, tid + T, ...).
template <class Body>
static void parallel_for(int n, const Body& body) {
    const int T = std::max(1, std::min({n, 16, (int)std::thread::hardware_concurrency()}));
    if (T <= 1) {
        for (int q = 0; q < n; ++q) body(q);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (int tid = 1; tid < T; ++tid)
        th.emplace_back([&, tid] { for (int q = tid; q < n; q += T) body(q); });
    for (int q = 0; q < n; q += T) body(q);
    for (auto& t : th) t.join();
}

static RegOps make_regops(const Flow& F, double beta) {
    RegOps R{};
    R.nx = F.nx; R.ny = F.ny; R.L = F.L; R.nb = F.nb;
    R.dx = F.dx; R.inv_dx = 1.0 / F.dx; R.bydx2 = 1. / (F.dx * F.dx); R.beta = beta;
    R.e_ptr = F.e_ptr; R.e_f = F.e_f; R.e_w = F.e_w;
    R.nnodes = F.nnodes; R.t_ptr = F.t_ptr; R.t_k = F.t_k; R.t_w = F.t_w;
    R.node_nbr = F.node_nbr; R.node_idx = F.node_idx;
    R.nold = F.nold; R.old_idx = F.old_idx;
    return R;
}

// Rebuilds the small per-flow regulariser descriptor arrays (node counts change
// when bodies move).
static int upload_regops(ibpm_batch* B) {
    const int nf = (int)B->flows.size();
    std::vector<RegOps> ic(nf), im(nf), cl(nf);
    B->max_nodes = 1;
    for (int k = 0; k < nf; ++k) {
        Flow& F = B->flows[k];
        ic[k] = make_regops(F, 0);
        ic[k].psi = F.psi; ic[k].b = F.b; ic[k].out = F.rhs;
        im[k] = make_regops(F, 0);
        im[k].psi = F.psi; im[k].b = nullptr; im[k].out = F.cg_q;
        im[k].skip = F.cg_state + 3;
        cl[k] = make_regops(F, 0);
        cl[k].cout = F.c;
        B->max_nodes = std::max(B->max_nodes, std::max(F.nnodes, F.nold));
    }
    CKR(up(B, B->d_interp_c, ic.data(), sizeof(RegOps) * nf));
    CKR(up(B, B->d_interp_m, im.data(), sizeof(RegOps) * nf));
    CKR(up(B, B->d_clear, cl.data(), sizeof(RegOps) * nf));
    for (int i = 0; i < B->nsub; ++i) {
        std::vector<RegOps> sf(nf), sd(nf);
        for (int k = 0; k < nf; ++k) {
            Flow& F = B->flows[k];
            sf[k] = make_regops(F, beta_of(B, F, i));
            sf[k].fin = F.f; sf[k].cout = F.c;
            sd[k] = make_regops(F, beta_of(B, F, i));
            sd[k].fin = F.cg_d; sd[k].cout = F.c;
            sd[k].skip = F.cg_state + 3;
        }
        CKR(up(B, B->d_smear_f[i], sf.data(), sizeof(RegOps) * nf));
        CKR(up(B, B->d_smear_d[i], sd.data(), sizeof(RegOps) * nf));
    }
    return 0;
}

static int alloc_flow(ibpm_batch* B, Flow& F) {
    DevMem& m = B->mem;
    const long ns = F.ns * F.L, nf = F.nf * F.L;
    CKR(m.alloc(&F.omega, ns)); CKR(m.alloc(&F.q, nf)); CKR(m.alloc(&F.nprev, ns));
    CKR(m.alloc(&F.a, ns)); CKR(m.alloc(&F.wstar, ns)); CKR(m.alloc(&F.psi, ns));
    CKR(m.alloc(&F.c, ns)); CKR(m.alloc(&F.U, ns)); CKR(m.alloc(&F.V, ns));
    CKR(m.alloc(&F.cross, nf)); CKR(m.alloc(&F.R, ns)); CKR(m.alloc(&F.T1, (size_t)(F.ny - 1) * t1_ld(F.nx - 1)));
    CKR(m.alloc(&F.T2, (size_t)(F.nx - 1) * t1_ld(F.ny - 1))); CKR(m.alloc(&F.mt, ns));
    CKR(cos_tables(m, F.nx, &F.cosx)); CKR(cos_tables(m, F.ny, &F.cosy));
    const int n = 2 * F.nb + 1;
    CKR(m.alloc(&F.f, n)); CKR(m.alloc(&F.rhs, n)); CKR(m.alloc(&F.b, n));
    CKR(m.alloc(&F.cg_r, n)); CKR(m.alloc(&F.cg_d, n)); CKR(m.alloc(&F.cg_q, n));
    CKR(m.alloc(&F.unit, n)); CKR(m.alloc(&F.mcol, (size_t)n * n));
    F.hist_cap = 1 << 16;
    CKR(m.alloc(&F.hist, 2 * (size_t)F.hist_cap)); CKR(m.alloc(&F.hist_count, 1));
    const int cap = 18 * F.nb + 1;
    CKR(m.alloc(&F.e_ptr, n + 1)); CKR(m.alloc(&F.e_f, cap)); CKR(m.alloc(&F.e_w, cap));
    CKR(m.alloc(&F.e_sep, (size_t)kESep * n));
    CKR(m.alloc(&F.t_ptr, cap + 1)); CKR(m.alloc(&F.t_k, cap)); CKR(m.alloc(&F.t_w, cap));
    CKR(m.alloc(&F.node_idx, 4 * cap)); CKR(m.alloc(&F.old_idx, 4 * cap)); CKR(m.alloc(&F.node_nbr, 4 * cap));
    {   // staging for the lists (capacities above), the constraint and the base flow
        const size_t bytes = sizeof(int) * (n + 1) + (sizeof(int) + sizeof(double)) * 2 * cap + sizeof(int) * (cap + 1) +
                             sizeof(int) * 8 * cap + sizeof(int4) * 4 * cap + sizeof(double) * (n + 4) +
                             sizeof(double) * kESep * n + 13 * 256;
        CK(hipHostMalloc((void**)&F.stage.p, bytes, hipHostMallocDefault));
        F.stage.cap = bytes;
    }
    if (F.p.model >= 1 && F.p.model <= 3) {   // base states
        const int np = F.p.model == 3 ? F.p.period : 1;
        F.bq.assign(np, nullptr);
        F.bw.assign(np, nullptr);
        F.base_set.assign(np, false);
        for (int k = 0; k < np; ++k) {
            CKR(m.alloc(&F.bq[k], nf));
            CKR(m.alloc(&F.bw[k], ns));
        }
    }
    if (F.p.model == 2) CKR(m.alloc(&F.sff, ns));
    if (F.p.model == 4) {
        CKR(m.alloc(&F.what, ns));
        CKR(m.alloc(&F.sfd_prev, ns));
    }
    {   // learned warm start history (small; allocated for every flow)
        const int R = 2 * B->nsub + 2;
        CKR(m.alloc(&F.ring_f, (size_t)R * n)); CKR(m.alloc(&F.ring_b, (size_t)R * n));
        CKR(m.alloc(&F.nn_x, (size_t)kNnIn * (F.nb + 1))); CKR(m.alloc(&F.nn_scale, 1));
    }
    return 0;
}

static SolveSpec spec_of(const Flow& F, int kind, double alpha, const double* f, double* u, double in_scale,
                         int epilogue, const double* W) {
    SolveSpec s{};
    s.nx = F.nx; s.ny = F.ny; s.L = F.L; s.dx = F.dx; s.kind = kind; s.alpha = alpha;
    s.f = f; s.u = u; s.R = F.R; s.T1 = F.T1; s.T2 = F.T2; s.in_scale = in_scale; s.epilogue = epilogue;
    s.W = W; s.cosx = F.cosx; s.cosy = F.cosy;
    return s;
}

static int build_plans(ibpm_batch* B) {
    Ctx* c = &B->ctx->c;
    const int nfl = (int)B->flows.size();
    CKR(B->mem.alloc(&B->cg_states, 8 * (size_t)nfl));
    for (int k = 0; k < nfl; ++k) B->flows[k].cg_state = B->cg_states + 8 * k;
    for (int i = 0; i < B->nsub; ++i) {
        std::vector<SolveSpec> s1, s2, s3, m1;
        for (auto& F : B->flows) {
            const double alpha = -beta_of(B, F, i) / 2. * (1. / F.p.reynolds);   // ProjectionSolver.cc:33
            s1.push_back(spec_of(F, 1, alpha, F.a, F.wstar, 1.0, 0, nullptr));
            // moving-body batches: omega = omega* too, for the local form of
            // omega = omega* - A^-1 c (plan_local_helm)
            if (B->any_cg) s1.back().u2 = F.omega;
            s2.push_back(spec_of(F, 0, 0.0, F.wstar, F.psi, -1.0, 0, nullptr));
            s3.push_back(spec_of(F, 1, alpha, F.c, F.omega, 1.0, 1, F.wstar));
            s3.back().vside = F.mt;   // M-temp is free after the force solve
            m1.push_back(spec_of(F, 1, alpha, F.c, F.mt, 1.0, 0, nullptr));
            m1.back().skip = F.cg_state + 3;
        }
        CKR(build_solve_plan(c, B->mem, s1, B->S1[i]));
        CKR(build_solve_plan(c, B->mem, s2, B->S2[i]));
        CKR(build_solve_plan(c, B->mem, s3, B->S3[i]));
        B->M1[i].force_solve = true;
        CKR(build_solve_plan(c, B->mem, m1, B->M1[i]));
    }
    std::vector<SolveSpec> s4, m2;
    for (auto& F : B->flows) {
        s4.push_back(spec_of(F, 0, 0.0, F.omega, F.psi, -1.0, 0, nullptr));
        m2.push_back(spec_of(F, 0, 0.0, F.mt, F.psi, -1.0, 0, nullptr));
        m2.back().skip = F.cg_state + 3;
    }
    CKR(build_solve_plan(c, B->mem, s4, B->S4));
    B->M2.force_solve = true;
    CKR(build_solve_plan(c, B->mem, m2, B->M2));
    // stencil descriptors
    for (int i = 0; i < B->nsub; ++i) {
        B->h_fops[i].clear();
        for (auto& F : B->flows) {
            FlowOps o{};
            o.nx = F.nx; o.ny = F.ny; o.L = F.L; o.dx = F.dx;
            o.q = F.q; o.omega = F.omega; o.U = F.U; o.V = F.V; o.cross = F.cross;
            o.a = F.a; o.nprev = F.nprev; o.psi = F.psi;
            const double nu = 1. / F.p.reynolds;   // NavierStokesModel::getAlpha
            o.lap_coef = 0.5 * nu * (B->an[i] + B->bn[i]);
            o.an = B->an[i]; o.bn = B->bn[i]; o.dt = F.p.dt; o.w_coef = 1.0;
            o.use_nprev = B->bn[i] != 0;
            o.nprev_valid = 0;
            const double pi = 4. * std::atan(1.);
            const double al = F.p.alpha_deg * pi / 180.;
            double sal, cal;
            sin_cos(al, sal, cal);
            o.ubase = F.base_mag * cal;
            o.vbase = F.base_mag * sal;
            o.ubf = F.ubf ? 1 : 0;
            o.bf = F.bf_dev;
            o.xc = F.bf.xc; o.yc = F.bf.yc;
            o.xoff = F.p.xoffset; o.yoff = F.p.yoffset;
            o.model = F.p.model;
            o.skip_central = F.p.model == 0;
            o.q0b = F.bq.empty() ? nullptr : F.bq[0];
            o.w0b = F.bw.empty() ? nullptr : F.bw[0];
            o.sff = F.sff;
            o.what = F.what;
            o.sfd_prev = F.sfd_prev;
            o.sfd_chi = F.p.sfd_chi;
            o.sfd_inv_delta = F.p.model == 4 ? 1.0 / F.p.sfd_delta : 0.0;   // Array.h:335-340
            B->h_fops[i].push_back(o);
        }
        CKR(B->mem.alloc(&B->d_fops[i], nfl));
        CKR(upload(B->d_fops[i], B->h_fops[i].data(), sizeof(FlowOps) * nfl));
        CKR(B->mem.alloc(&B->d_smear_f[i], nfl));
        CKR(B->mem.alloc(&B->d_smear_d[i], nfl));
        CKR(B->mem.alloc(&B->d_minv[i], nfl));
    }
    CKR(B->mem.alloc(&B->d_interp_c, nfl));
    CKR(B->mem.alloc(&B->d_interp_m, nfl));
    CKR(B->mem.alloc(&B->d_clear, nfl));
    std::vector<CgOps> cg(nfl);
    std::vector<ForceOps> fo(nfl);
    for (int k = 0; k < nfl; ++k) {
        Flow& F = B->flows[k];
        cg[k] = CgOps{2 * F.nb, F.f, F.cg_r, F.cg_d, F.cg_q, F.rhs, F.cg_state};
        fo[k] = ForceOps{F.nb, F.f, F.hist, F.hist_count};
    }
    CKR(B->mem.alloc(&B->d_cg, nfl));
    CKR(upload(B->d_cg, cg.data(), sizeof(CgOps) * nfl));
    CKR(B->mem.alloc(&B->d_force, nfl));
    CKR(upload(B->d_force, fo.data(), sizeof(ForceOps) * nfl));
    return 0;
}

static void mark_nprev_valid(ibpm_batch* B) {
    if (B->old_saved) return;
    B->old_saved = true;
    for (int i = 0; i < B->nsub; ++i) {
        for (auto& o : B->h_fops[i]) o.nprev_valid = o.sfd_valid = 1;
        (void)hipMemcpy(B->d_fops[i], B->h_fops[i].data(), sizeof(FlowOps) * B->h_fops[i].size(),
                        hipMemcpyHostToDevice);
    }
}

// The local Helmholtz path of substep i's M applications (k_fs_helm).  The
// right-hand side of A^-1 inside M is beta B f: nonzero only at the
// regulariser's nodes on level 0, and on coarser levels where the coarsify
// chain carries it (coarse point ic sees fine rows 2 (ic - ex) - 1 .. + 1).
// Per flow and level the box is that support grown by the sweep count S,
// chosen so that rho^S <= 1e-18, rho = 4|c| / |1 - 4c| the Jacobi
// contraction.  The path is off for the substep (the DST solve runs) when any
// box would reach a level's first or last interior row or column, where
// getBC's values enter, or a coarser level's ring, where the finer level
// reads them; the solution there is then below 1e-18 of its maximum.
constexpr int kHelmMaxSweeps = 40;

// The L^-1 solve that follows the local A^-1 runs its passes over part of
// each level.  Pass 3 writes only the psi that is read afterwards: on level
// 0, the rows of the regulariser's nodes (and one more on each side), where
// the interpolation takes Curl psi; on coarser levels, the ring that the next
// finer level's getBC reads (rows ex and ex + nx/2 whole, columns ey and
// ey + ny/2 of the rows between).  The solve's other consumers are none: the
// step's refresh solves psi again.  Pass 1 takes rows off the right-hand
// side's support in closed form: the support is the local result's box on
// each level, with coarsify's reach from the finer level inside the central
// region.  (The LDS passes ignore both restrictions and do the whole level.)
// Algorithmic bytes of each pass of a job under its restrictions (EllJob's
// partial-pass fields): 8 B per point read and per point written, counting
// only the rows a pass reads and writes.
static void pass_bytes(const EllJob& J, const PassGroup& G, double out[3]) {
    const double N1 = J.nx - 1, N2 = J.ny - 1;
    if (G.fp32) {   // whole level: FP64 in, FP32 T1; FP32 T1 -> T2; FP32 T2, FP64 out
        out[0] = 12.0 * N1 * N2;
        out[1] = 8.0 * N1 * N2;
        out[2] = 12.0 * N1 * N2;
        return;
    }
    const int kG1 = r16_pass_group(1);   // row pairs per workgroup of the register-held pass 1
    // pass 1: T1 whole; rows read unless their workgroup takes the closed form
    double read1 = N1;
    if (J.sin1 && G.regs[0]) {
        read1 = 0;
        for (int r0 = 0; r0 < (int)N1; r0 += 2 * kG1) {
            const int i_first = r0 + 1, i_last = std::min(r0 + 2 * kG1, (int)N1);
            const bool closed = (i_last < J.sup_lo || i_first > J.sup_hi) && i_first > 1 && i_last < J.nx - 1;
            if (!closed) read1 += i_last - i_first + 1;
        }
    }
    out[0] = 8.0 * N2 * (read1 + N1);
    // pass 2: T1 whole in, T2 rows [t2_lo, t2_hi] out
    const double w2 = std::max(0.0, std::min((double)J.t2_hi, N1 - 1) - std::max(J.t2_lo, 0) + 1);
    out[1] = 8.0 * N2 * (N1 + w2);
    // pass 3: rows of transformed pairs in and out, ring rows in and two values out
    double full = 0, ring = 0;
    for (int q = 0; 2 * q < (int)N1; ++q) {
        const int rows = std::min(2, (int)N1 - 2 * q);
        const bool f = (q >= J.full_lo && q <= J.full_hi) || q == J.fq0 || q == J.fq1;
        if (f) {
            full += rows;
        } else {
            for (int h = 0; h < rows; ++h) {
                const int i = 2 * q + h + 1;
                if (i >= J.ring_lo && i <= J.ring_hi) ring += 1;
            }
        }
    }
    out[2] = 8.0 * (N2 * (2 * full + ring) + 2 * ring);
}

// Per-pass bytes of a group's jobs after their restrictions changed.
static void group_bytes(PassGroup& G) {
    for (int q = 0; q < 3; ++q) G.bytes[q] = 0;
    for (size_t j = 0; j < G.h.size(); ++j) {
        double b[3];
        pass_bytes(G.h[j], G, b);
        for (int q = 0; q < 3; ++q) {
            G.jbytes[q][j] = b[q];
            G.bytes[q] += b[q];
        }
    }
}

// Passes 3 and 2 of a Poisson solve whose psi is read only by the
// interpolation (at the regulariser's nodes, level-0 rows node_lo..node_hi)
// and by the finer levels' getBC (the coarser levels' rings): see
// restrict_m2l_pass3.  Register-held pass 3 only.
static void read_rows_only(EllJob& J, const PassGroup& G, int node_lo, int node_hi, const double* ringsin) {
    if (!G.regs[2] || J.epilogue != 0 || J.vside) return;
    const int N1 = J.nx - 1, ex = nx_ext(J.nx), ey = nx_ext(J.ny);
    if (J.lev == 0) {
        J.full_lo = std::max(node_lo - 2, 0) / 2;
        J.full_hi = std::min(node_hi, N1 - 1) / 2;
        J.t2_lo = 2 * J.full_lo;
        J.t2_hi = 2 * J.full_hi + 1;
    } else {
        J.full_lo = 1; J.full_hi = 0;
        J.fq0 = (ex - 1) / 2; J.fq1 = (ex + J.nx / 2 - 1) / 2;
        J.ring_lo = ex + 1; J.ring_hi = ex + J.nx / 2 - 1;
        J.rj0 = ey; J.rj1 = ey + J.ny / 2;
        J.ringsin = ringsin;
        J.t2_lo = 2 * J.fq0;
        J.t2_hi = 2 * J.fq1 + 1;
    }
    if (!G.regs[1]) {   // an LDS pass 2 writes every row
        J.t2_lo = 0;
        J.t2_hi = INT_MAX;
    }
}

static int ring_tables(ibpm_batch* B) {
    for (auto& F : B->flows) {
        if (F.ringsin) continue;
        const int n = F.ny - 1, ey = nx_ext(F.ny);
        std::vector<double> t(2 * (size_t)n), s1(n);
        const double pi = 4. * std::atan(1.);
        for (int l = 1; l <= n; ++l) {
            t[l - 1] = 2. * std::sin(pi * l * ey / F.ny);
            t[n + l - 1] = 2. * std::sin(pi * l * (ey + F.ny / 2) / F.ny);
            s1[l - 1] = 2. * std::sin(pi * l / F.ny);
        }
        CKR(B->mem.alloc(&F.ringsin, t.size()));
        CKR(upload(F.ringsin, t.data(), sizeof(double) * t.size()));
        CKR(B->mem.alloc(&F.sin1, s1.size()));
        CKR(upload(F.sin1, s1.data(), sizeof(double) * s1.size()));
    }
    return 0;
}

// The step's psi = L^-1(-omega*) of substep i feeds only the interpolation of
// C omega* (the force solve overwrites psi and the refresh solves it again),
// so its passes 3 and 2 take read_rows_only from the current node rows.
static int restrict_s2(ibpm_batch* B, int i) {
    CKR(ring_tables(B));
    std::vector<std::array<int, 2>> rows(B->flows.size(), std::array<int, 2>{1, INT_MAX});
    for (size_t k = 0; k < B->flows.size(); ++k) {
        const Flow& F = B->flows[k];
        if (F.reg.node_idx.empty()) continue;
        int lo = INT_MAX, hi = INT_MIN;
        for (int idx : F.reg.node_idx) {
            lo = std::min(lo, idx / (F.ny - 1) + 1);
            hi = std::max(hi, idx / (F.ny - 1) + 1);
        }
        rows[k] = {lo, hi};
    }
    for (auto& step : B->S2[i].steps)
        for (auto& G : step) {
            bool changed = false;
            for (size_t j = 0; j < G.h.size(); ++j) {
                EllJob& J = G.h[j];
                const EllJob before = J;
                whole_level(J);
                const int k = G.jflow[j];
                if (B->helm_allowed && rows[k][1] != INT_MAX)
                    read_rows_only(J, G, rows[k][0], rows[k][1], B->flows[k].ringsin);
                if (std::memcmp(&before, &J, sizeof(EllJob)) != 0) changed = true;
            }
            if (changed) {
                CKR(up(B, G.d, G.h.data(), sizeof(EllJob) * G.h.size()));
                group_bytes(G);
            }
        }
    return 0;
}

static int restrict_m2l_pass3(ibpm_batch* B, const std::vector<std::vector<std::array<int, 4>>>& nsup) {
    CKR(ring_tables(B));
    // per flow and level: rbox, where the coarsify chain puts the right-hand
    // side's R (coarsify's reach from the finer level's support, inside the
    // central region; empty on level 0), and sup, the bounding box of the
    // right-hand side (rbox and the local result's box)
    const int nfl = (int)B->flows.size();
    std::vector<std::vector<std::array<int, 4>>> rbox(nfl), sup(nfl);
    for (int k = 0; k < nfl; ++k) {
        const Flow& F = B->flows[k];
        const int ex = nx_ext(F.nx), ey = nx_ext(F.ny);
        rbox[k].resize(F.L);
        sup[k].resize(F.L);
        rbox[k][0] = {1, 0, 1, 0};
        sup[k][0] = F.hbox[0];
        for (int lev = 1; lev < F.L; ++lev) {
            const std::array<int, 4>& fs = sup[k][lev - 1];
            std::array<int, 4> r = {std::max(fs[0] / 2 + ex, ex + 1), std::min((fs[1] + 1) / 2 + ex, ex + F.nx / 2 - 1),
                                    std::max(fs[2] / 2 + ey, ey + 1), std::min((fs[3] + 1) / 2 + ey, ey + F.ny / 2 - 1)};
            rbox[k][lev] = r;
            const std::array<int, 4>& h = F.hbox[lev];
            sup[k][lev] = (r[0] <= r[1] && r[2] <= r[3])
                              ? std::array<int, 4>{std::min(r[0], h[0]), std::max(r[1], h[1]), std::min(r[2], h[2]),
                                                   std::max(r[3], h[3])}
                              : h;
        }
    }
    // the boxed coarsify chain needs every pass 1 of the solve to mask R
    bool all_regs1 = true;
    for (auto& step : B->M2L.steps)
        for (auto& G : step) all_regs1 = all_regs1 && G.regs[0];
    B->m2l_boxed = all_regs1;
    for (auto& step : B->M2L.steps)
        for (auto& G : step) {
            bool changed = false;
            for (size_t j = 0; j < G.h.size(); ++j) {
                EllJob& J = G.h[j];
                const int k = G.jflow[j];
                const EllJob before = J;
                whole_level(J);
                if (J.epilogue == 0 && !J.vside) {
                    J.sup_lo = sup[k][J.lev][0];
                    J.sup_hi = sup[k][J.lev][1];
                    J.sin1 = B->flows[k].sin1;
                    if (B->m2l_boxed) {
                        const std::array<int, 4>& r = rbox[k][J.lev];
                        J.rb_i0 = r[0]; J.rb_i1 = r[1]; J.rb_j0 = r[2]; J.rb_j1 = r[3];
                    }
                }
                read_rows_only(J, G, nsup[k][0][0], nsup[k][0][1], B->flows[k].ringsin);
                if (std::memcmp(&before, &J, sizeof(EllJob)) != 0) changed = true;
            }
            if (changed) {
                CKR(up(B, G.d, G.h.data(), sizeof(EllJob) * G.h.size()));
                group_bytes(G);
            }
        }
    if (!B->m2l_boxed) return 0;
    std::vector<HelmJob> cj;
    B->m2lc_flow.clear();
    B->m2lc_bytes.clear();
    long maxa = 1;
    for (int k = 0; k < nfl; ++k) {
        const Flow& F = B->flows[k];
        for (int lev = 1; lev < F.L; ++lev) {
            const std::array<int, 4>& r = rbox[k][lev];
            const std::array<int, 4>& fr = rbox[k][lev - 1];
            HelmJob J{};
            J.nx = F.nx; J.ny = F.ny; J.lev = lev;
            J.s0 = r[0]; J.s1 = r[1]; J.t0 = r[2]; J.t1 = r[3];
            J.fs0 = fr[0]; J.fs1 = fr[1]; J.ft0 = fr[2]; J.ft1 = fr[3];
            J.f_fine = F.m1 + (lev - 1) * F.ns; J.R_fine = F.R + (lev - 1) * F.ns;
            J.R = F.R + lev * F.ns;
            J.skip = F.cg_state + 3;
            const long area = (r[0] <= r[1] && r[2] <= r[3]) ? (long)(r[1] - r[0] + 1) * (r[3] - r[2] + 1) : 0;
            maxa = std::max(maxa, area);
            cj.push_back(J);
            B->m2lc_flow.push_back(k);
            B->m2lc_bytes.push_back(40.0 * area);
        }
    }
    if ((int)cj.size() > B->m2lc_cap) {
        CKR(B->mem.alloc(&B->d_m2lc, cj.size()));
        B->m2lc_cap = (int)cj.size();
    }
    CKR(up(B, B->d_m2lc, cj.data(), sizeof(HelmJob) * cj.size()));
    B->m2lc_n = (int)cj.size();
    B->m2lc_blocks = (int)((maxa + kThreads - 1) / kThreads);
    return 0;
}

static int plan_local_helm(ibpm_batch* B, int i) {
    B->helm_on[i] = false;
    if (!B->helm_allowed) return 0;
    const int nfl = (int)B->flows.size();
    std::vector<std::vector<std::array<int, 4>>> nbox(nfl), nsup(nfl);
    int sweeps = 1;
    for (int k = 0; k < nfl; ++k) {
        Flow& F = B->flows[k];
        const std::vector<int>& nodes = F.reg.node_idx;
        if (nodes.empty()) return 0;
        const int ny1 = F.ny - 1;
        int s0 = INT_MAX, s1 = INT_MIN, t0 = INT_MAX, t1 = INT_MIN;   // support: rows i, columns j (1-based)
        for (int idx : nodes) {
            const int ii = idx / ny1 + 1, jj = idx % ny1 + 1;
            s0 = std::min(s0, ii); s1 = std::max(s1, ii);
            t0 = std::min(t0, jj); t1 = std::max(t1, jj);
        }
        const double alpha = -beta_of(B, F, i) / 2. * (1. / F.p.reynolds);   // as the M1 solve's
        const int ex = nx_ext(F.nx), ey = nx_ext(F.ny);
        nbox[k].resize(F.L);
        nsup[k].resize(F.L);
        for (int lev = 0; lev < F.L; ++lev) {
            if (lev >= 1) {   // support after coarsify, inside the central region
                s0 = std::max((s0 - 1 + 1) / 2 + ex, ex + 1);   // ceil((s0 - 1) / 2), s0 >= 1
                s1 = std::min((s1 + 1) / 2 + ex, ex + F.nx / 2 - 1);
                t0 = std::max((t0 - 1 + 1) / 2 + ey, ey + 1);
                t1 = std::min((t1 + 1) / 2 + ey, ey + F.ny / 2 - 1);
                if (s0 > s1 || t0 > t1) return 0;
            }
            const double dxl = F.dx * (1 << lev), c = alpha / (dxl * dxl);
            const double rho = std::fabs(4 * c) / std::fabs(1 - 4 * c);
            if (!(rho < 0.5)) return 0;
            const int S = rho < 1e-300 ? 1 : (int)std::ceil(std::log(1e-18) / std::log(rho));
            if (S > kHelmMaxSweeps) return 0;
            std::array<int, 4> b = {s0 - S, s1 + S, t0 - S, t1 + S};
            if (b[0] < 2 || b[1] > F.nx - 2 || b[2] < 2 || b[3] > F.ny - 2) return 0;
            if (lev >= 1 && !(b[0] > ex && b[1] < ex + F.nx / 2 && b[2] > ey && b[3] < ey + F.ny / 2)) return 0;
            nbox[k][lev] = b;
            nsup[k][lev] = {s0, s1, t0, t1};
            sweeps = std::max(sweeps, S);
        }
    }
    std::vector<HelmJob> jobs, jobs3, clears;
    std::vector<int> jflow;
    std::vector<double> jbytes, cbytes;
    long maxarea = 0, maxsup = 0, maxpts = 1;
    for (int k = 0; k < nfl; ++k) {
        Flow& F = B->flows[k];
        const long ns = F.ns * F.L;
        if (!F.m1) {
            CKR(B->mem.alloc(&F.m1, ns));
            CKR(B->mem.alloc(&F.m1tmp, ns));
        }
        // the result and sweep buffers are zero off the current boxes: clear
        // what an earlier box covered (the result then depends on the current
        // boxes alone, not on the run's history)
        for (int lev = 0; lev < (int)F.hbox.size() && lev < F.L; ++lev)
            if (F.hbox[lev] != nbox[k][lev]) {
                HelmJob C{};
                C.nx = F.nx; C.ny = F.ny; C.lev = lev;
                C.i0 = F.hbox[lev][0]; C.i1 = F.hbox[lev][1]; C.j0 = F.hbox[lev][2]; C.j1 = F.hbox[lev][3];
                C.out = F.m1 + lev * F.ns; C.tmp = F.m1tmp + lev * F.ns;
                clears.push_back(C);
            }
        F.hbox = nbox[k];
        const double alpha = -beta_of(B, F, i) / 2. * (1. / F.p.reynolds);
        for (int lev = 0; lev < F.L; ++lev) {
            const std::array<int, 4>& b = nbox[k][lev];
            const std::array<int, 4>& sp = nsup[k][lev];
            const double dxl = F.dx * (1 << lev), c = alpha / (dxl * dxl);
            HelmJob J{};
            J.nx = F.nx; J.ny = F.ny; J.lev = lev;
            J.i0 = b[0]; J.i1 = b[1]; J.j0 = b[2]; J.j1 = b[3];
            J.s0 = sp[0]; J.s1 = sp[1]; J.t0 = sp[2]; J.t1 = sp[3];
            if (lev >= 1) {
                const std::array<int, 4>& fs = nsup[k][lev - 1];
                J.fs0 = fs[0]; J.fs1 = fs[1]; J.ft0 = fs[2]; J.ft1 = fs[3];
                J.f_fine = F.c + (lev - 1) * F.ns; J.R_fine = F.R + (lev - 1) * F.ns;
                const long sa = (long)(sp[1] - sp[0] + 1) * (sp[3] - sp[2] + 1);
                maxsup = std::max(maxsup, sa);
                cbytes.push_back(40.0 * sa);   // 9 fine reads (4 new per point) and one write
            } else {
                cbytes.push_back(0.0);
            }
            J.c = c; J.inv_diag = 1. / (1 - 4 * c); J.in_scale = 1.0;
            J.f = F.c + lev * F.ns; J.R = F.R + lev * F.ns;
            J.out = F.m1 + lev * F.ns; J.tmp = F.m1tmp + lev * F.ns;
            J.skip = F.cg_state + 3;
            const long area = (long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1);
            maxarea = std::max(maxarea, (long)helm_tiles(b[1] - b[0] + 1, b[3] - b[2] + 1));
            maxpts = std::max(maxpts, area);
            jobs.push_back(J);
            HelmJob J3 = J;   // omega = omega* - A^-1 c after the force solve: every flow, always
            J3.skip = nullptr;
            J3.eo = F.omega + lev * F.ns;
            J3.ei = F.wstar + lev * F.ns;
            jobs3.push_back(J3);
            jflow.push_back(k);
            jbytes.push_back(24.0 * area * std::max(1, sweeps));   // per sweep: r and u in, u out
        }
    }
    if (!B->m2l_built) {   // L^-1 of the local result (the DST solve M2 reads mt instead)
        std::vector<SolveSpec> m2l;
        for (auto& F : B->flows) {
            m2l.push_back(spec_of(F, 0, 0.0, F.m1, F.psi, -1.0, 0, nullptr));
            m2l.back().skip = F.cg_state + 3;
        }
        B->M2L.force_solve = true;
        CKR(build_solve_plan(&B->ctx->c, B->mem, m2l, B->M2L));
        if (B->fs32) set_plan_fp32(B->M2L, true);
        B->m2l_built = true;
    }
    CKR(restrict_m2l_pass3(B, nsup));
    if ((int)jobs.size() > B->helm_cap) {
        CKR(B->mem.alloc(&B->d_helm, jobs.size()));
        CKR(B->mem.alloc(&B->d_helm3, jobs.size()));
        B->helm_cap = (int)jobs.size();
    }
    CKR(up(B, B->d_helm, jobs.data(), sizeof(HelmJob) * jobs.size()));
    CKR(up(B, B->d_helm3, jobs3.data(), sizeof(HelmJob) * jobs3.size()));
    {
        std::vector<HelmJob> j0;
        long t0 = 1;
        B->helm0_bytes = 0;
        for (size_t q = 0; q < jobs.size(); ++q)
            if (jobs[q].lev == 0) {
                j0.push_back(jobs[q]);
                t0 = std::max(t0, (long)helm_tiles(jobs[q].i1 - jobs[q].i0 + 1, jobs[q].j1 - jobs[q].j0 + 1));
                B->helm0_bytes += jbytes[q];
            }
        if (!B->d_helm0) CKR(B->mem.alloc(&B->d_helm0, (size_t)nfl));
        CKR(up(B, B->d_helm0, j0.data(), sizeof(HelmJob) * j0.size()));
        B->helm0_n = (int)j0.size();
        B->helm0_blocks = (int)t0;
    }
    if (!clears.empty()) {
        long maxc = 1;
        for (const HelmJob& C : clears) maxc = std::max(maxc, (long)(C.i1 - C.i0 + 1) * (C.j1 - C.j0 + 1));
        if ((int)clears.size() > B->hclear_cap) {
            CKR(B->mem.alloc(&B->d_hclear, clears.size()));
            B->hclear_cap = (int)clears.size();
        }
        CKR(up(B, B->d_hclear, clears.data(), sizeof(HelmJob) * clears.size()));
        if (B->async_up) CKR(async_flush(B));   // the pieces in place before the launch that reads them
        launch_helm_clear(B->d_hclear, (int)clears.size(), (int)((maxc + kThreads - 1) / kThreads),
                          B->ctx->c.stream);
    }
    B->helm_ablocks = (int)((maxpts + kThreads - 1) / kThreads);
    B->helm_n = (int)jobs.size();
    B->helm_blocks = (int)maxarea;   // tiles of the largest box
    B->helm_cblocks = (int)((maxsup + kThreads - 1) / kThreads);
    B->helm_sweeps = sweeps;
    B->helm_flow = jflow;
    B->helm_bytes = jbytes;
    B->helm_cbytes = cbytes;
    B->helm_on[i] = true;
    return 0;
}

// Algorithmic bytes of a k_interp (with_b: also the constraint b, 8 B per
// component) or k_smear_curl launch over the flows of `active` (null: all).
static double reg_bytes(const ibpm_batch* B, bool smear, bool with_b, const std::vector<char>* active = nullptr) {
    double by = 0;
    for (size_t k = 0; k < B->flows.size(); ++k) {
        if (active && !(*active)[k]) continue;
        const Flow& F = B->flows[k];
        by += smear ? F.by_smear : F.by_interp + (with_b ? 16.0 * F.nb : 0.0);
    }
    return by;
}

// M f = C A^-1 (beta B f) for the vector behind d_smear (ProjectionSolver.cc:96-101).
// interp: the interpolation's descriptors (null: B->d_interp_m); single_level:
// L^-1 on the finest level alone (plan M1L: the preconditioner's
// approximate M), local A^-1 path only.
static void enqueue_M(ibpm_batch* B, int i, RegOps* smear, hipStream_t s,
                      const std::vector<char>* active = nullptr, const RegOps* interp = nullptr,
                      bool single_level = false) {
    const int nfl = (int)B->flows.size();
    Probe* pr = B->probe;
    timed(pr, K_SMEAR, pr ? reg_bytes(B, true, false, active) : 0, s,
          [&] { launch_smear_curl(smear, nfl, B->max_nodes, s); });
    if (B->helm_on[i]) {   // A^-1 by local sweeps (plan_local_helm), then L^-1 of their result
        double by = 0, cby = 0;
        if (pr)
            for (size_t j = 0; j < B->helm_flow.size(); ++j)
                if (!active || (*active)[B->helm_flow[j]]) {
                    by += B->helm_bytes[j];
                    cby += B->helm_cbytes[j];
                }
        for (int lev = 1; lev < (single_level ? 1 : B->M1[i].nlev); ++lev)
            timed(pr, K_FSCOARSIFY, cby / std::max(1, B->M1[i].nlev - 1), s,
                  [&] { launch_helm_coarsify(B->d_helm, B->helm_n, B->helm_cblocks, lev, s); });
        const int nl = helm_launches(B->helm_sweeps);
        if (single_level)   // level 0's sweeps alone: the coarser levels' results are not read
            launch_helm_sweeps(B->d_helm0, B->helm0_n, B->helm0_blocks, B->helm_sweeps, s,
                               [&](const std::function<void()>& fn) { timed(pr, K_FSHELM, B->helm0_bytes / nl, s, fn); });
        else
            launch_helm_sweeps(B->d_helm, B->helm_n, B->helm_blocks, B->helm_sweeps, s,
                               [&](const std::function<void()>& fn) { timed(pr, K_FSHELM, by / nl, s, fn); });
        const bool boxed = B->m2l_boxed && !B->M2L.fp32;   // the boxed chain feeds register-held passes only
        if (single_level) {
            run_solve(B->M1L, s, pr, active, false);
        } else if (boxed) {
            double cb = 0;
            if (pr)
                for (size_t j = 0; j < B->m2lc_flow.size(); ++j)
                    if (!active || (*active)[B->m2lc_flow[j]]) cb += B->m2lc_bytes[j];
            for (int lev = 1; lev < B->M2L.nlev; ++lev)
                timed(pr, K_FSCOARSIFY, cb / std::max(1, B->M2L.nlev - 1), s,
                      [&] { launch_helm_coarsify(B->d_m2lc, B->m2lc_n, B->m2lc_blocks, lev, s); });
        }
        if (!single_level) run_solve(B->M2L, s, pr, active, !boxed);
        B->st_local_mapps += 1;
    } else {
        run_solve(B->M1[i], s, pr, active);
        run_solve(B->M2, s, pr, active);
    }
    timed(pr, K_INTERP, pr ? reg_bytes(B, false, false, active) : 0, s,
          [&] { launch_interp(interp ? interp : B->d_interp_m, nfl, B->max_comp, s); });
}

// Nonlinear term and right-hand side a of substep i (IBSolver.cc:144,169-183).
static void enqueue_rhs(ibpm_batch* B, int i, hipStream_t s) {
    const int nfl = (int)B->flows.size();
    int lmax = 0;
    for (auto& F : B->flows) lmax = std::max(lmax, F.L);
    Probe* pr = B->probe;
    if (B->cff_n) {   // adjoint flows: s = CrossProduct(q0, q), coarsified (VectorOperations.cc:465-486)
        launch_cross_ff(B->d_cff, B->cff_n, B->max_ns_all, s);
        for (size_t lev = 1; lev < B->d_cffc.size(); ++lev)
            if (B->cffc_n[lev]) launch_coarsify(B->d_cffc[lev], B->cffc_n[lev], B->cffc_rows, s);
    }
    timed(pr, K_VEL, B->bytes_vel, s, [&] { launch_vel(B->d_fops[i], nfl, B->max_ns_all, s); });
    for (int lev = 0; lev < lmax; ++lev)
        timed(pr, K_CROSS, B->bytes_cross / lmax, s, [&] { launch_cross(B->d_fops[i], nfl, B->max_nf, lev, s); });
    timed(pr, K_RHS, B->bytes_rhs, s, [&] { launch_rhs(B->d_fops[i], nfl, B->max_ns_all, s); });
}

static void enqueue_refresh(ibpm_batch* B, hipStream_t s) {
    run_solve(B->S4, s, B->probe);
    timed(B->probe, K_CURLQ, B->bytes_curlq, s,
          [&] { launch_curl_q(B->d_fops[0], (int)B->flows.size(), B->max_cells_all, s); });
}

// The stationary-body (Cholesky) substep: fixed launch sequence.
static void enqueue_substep_chol(ibpm_batch* B, int i, hipStream_t s) {
    const int nfl = (int)B->flows.size();
    Probe* pr = B->probe;
    enqueue_rhs(B, i, s);
    run_solve(B->S1[i], s, pr);                     // omega* = A^-1 a
    run_solve(B->S2[i], s, pr);                     // psi = L^-1(-omega*)
    timed(pr, K_INTERP, pr ? reg_bytes(B, false, true) : 0, s,
          [&] { launch_interp(B->d_interp_c, nfl, B->max_comp, s); });   // rhs = C omega* - b
    double by_minv = 0;   // the n x n inverse, the right-hand side and the output, FP64
    if (pr)
        for (auto& F : B->flows) by_minv += 8.0 * (4.0 * F.nb * F.nb + 4.0 * F.nb);
    timed(pr, K_MINV, by_minv, s, [&] { launch_minv(B->d_minv[i], nfl, B->max_n, s); });   // f = M^-1 rhs
    timed(pr, K_SMEAR, pr ? reg_bytes(B, true, false) : 0, s,
          [&] { launch_smear_curl(B->d_smear_f[i], nfl, B->max_nodes, s); });   // c = beta B f
    run_solve(B->S3[i], s, pr);                     // omega = omega* - A^-1 c
    enqueue_refresh(B, s);                          // q = Curl psi + q0
}

static int enqueue_step_chol(ibpm_batch* B, hipStream_t s) {
    for (int i = 0; i < B->nsub; ++i) enqueue_substep_chol(B, i, s);
    timed(B->probe, K_FORCES, 0, s, [&] { launch_forces(B->d_force, (int)B->flows.size(), s); });
    return 0;
}

// Descriptor of the learned warm start for substep k = B->ksub (target: the
// final force of the substep, for sample collection).
// The descriptors of every flow for substep k, RK substep i (ring slots
// relative to k, whose position in the ring alone matters).
static void fill_nn(const ibpm_batch* B, long k, int i, NnOps* ops) {
    const int S = B->nsub, R = B->ring;
    auto slot = [&](const Flow& F, long j) { return F.ring_f + (size_t)(((j % R) + R) % R) * (2 * F.nb + 1); };
    auto bslot = [&](const Flow& F, long j) { return F.ring_b + (size_t)(((j % R) + R) % R) * (2 * F.nb + 1); };
    for (size_t q = 0; q < B->flows.size(); ++q) {
        const Flow& F = B->flows[q];
        NnOps& o = ops[q];
        o = NnOps{};
        o.nb = F.nb; o.substep = i;
        o.fa = slot(F, k - 1); o.fb = slot(F, k - S); o.fc = slot(F, k - S - 1);
        o.fd = slot(F, k - 2 * S); o.fe = slot(F, k - 2 * S - 1);
        o.b_now = F.b; o.b_prev = bslot(F, k - 1);
        o.X = F.nn_x; o.scale = F.nn_scale; o.params = B->nn_params; o.x0 = F.f;
    }
}

static int upload_nn(ibpm_batch* B, int i, bool with_target) {
    const int nfl = (int)B->flows.size();
    std::vector<NnOps> ops(nfl);
    fill_nn(B, B->ksub, i, ops.data());
    long off = B->data_n;
    if (with_target)
        for (int q = 0; q < nfl; ++q) {
            NnOps& o = ops[q];
            o.target = B->flows[q].f; o.data_x = B->data_x; o.data_y = B->data_y; o.data_off = off;
            off += B->flows[q].nb;
        }
    CKR(upload(B->d_nn, ops.data(), sizeof(NnOps) * nfl));
    if (with_target) B->data_n = off;
    return 0;
}

// Inference descriptors of every (ring position, RK substep), uploaded once,
// so that the learned warm start launches without a host round trip:
// entry (k mod ring) nsub + i for substep k.
static int nn_infer_table(ibpm_batch* B) {
    if (B->d_nn_inf) return 0;
    const int nfl = (int)B->flows.size();
    std::vector<NnOps> all((size_t)B->ring * B->nsub * nfl);
    for (int p = 0; p < B->ring; ++p)
        for (int i = 0; i < B->nsub; ++i) fill_nn(B, p, i, all.data() + ((size_t)p * B->nsub + i) * nfl);
    CKR(B->mem.alloc(&B->d_nn_inf, all.size()));
    CKR(upload(B->d_nn_inf, all.data(), sizeof(NnOps) * all.size()));
    return 0;
}

static bool nn_history_ready(const ibpm_batch* B) { return B->ksub >= 2L * B->nsub + 1; }

static int nn_samples_room(const ibpm_batch* B) {
    long need = 0;
    for (auto& F : B->flows) need += F.nb;
    return B->data_n + need <= B->data_cap;
}

static int build_precond(ibpm_batch* B, int i);
static int refresh_precond(ibpm_batch* B, int i);
static int join_refresh(ibpm_batch* B);

// Force-solve state of every flow: zero, with tol^2 in slot 1.
// Pinned host buffers and the event of the solver polls (first use).
static int poll_buffers(ibpm_batch* B) {
    if (B->h_states) return 0;
    const size_t n = 8 * B->flows.size();
    CK(hipHostMalloc((void**)&B->h_states, sizeof(double) * n, hipHostMallocDefault));
    CK(hipHostMalloc((void**)&B->h_reset, sizeof(double) * n, hipHostMallocDefault));
    CK(hipEventCreateWithFlags(&B->poll_ev, hipEventDisableTiming));
    return 0;
}

// Queued on the stream (no wait): the states are read by the kernels after
// it, and the pinned source is rewritten only by the next solve, after polls
// that have waited past this copy.
static int reset_states(ibpm_batch* B) {
    hipStream_t s = B->ctx->c.stream;
    const int nfl = (int)B->flows.size();
    CKR(poll_buffers(B));
    for (int k = 0; k < 8 * nfl; ++k) B->h_reset[k] = 0.0;
    for (int k = 0; k < nfl; ++k) B->h_reset[8 * k + 1] = B->flows[k].p.cg_tol * B->flows[k].p.cg_tol;
    CK(hipMemcpyAsync(B->cg_states, B->h_reset, sizeof(double) * 8 * nfl, hipMemcpyHostToDevice, s));
    return 0;
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A poll in two halves: the states copied to pinned memory right behind the
// work queued so far (poll_begin), so that the GPU may be handed more work
// before the host waits for the copy (poll_wait).
static int poll_begin(ibpm_batch* B) {
    hipStream_t s = B->ctx->c.stream;
    CKR(poll_buffers(B));
    CK(hipMemcpyAsync(B->h_states, B->cg_states, sizeof(double) * 8 * B->flows.size(), hipMemcpyDeviceToHost, s));
    CK(hipEventRecord(B->poll_ev, s));
    return 0;
}

static int poll_wait(ibpm_batch* B, std::vector<double>& st) {
    const double t0 = now_s();
    CK(hipEventSynchronize(B->poll_ev));
    st.assign(B->h_states, B->h_states + 8 * B->flows.size());
    B->st_host_poll_s += now_s() - t0;
    return 0;
}

static int read_states(ibpm_batch* B, std::vector<double>& st) {
    CKR(poll_begin(B));
    return poll_wait(B, st);
}

// Zero-fill pieces of every flow's f, then of its cg_q (first use).
static int zero_pieces(ibpm_batch* B) {
    if (B->d_zero_pieces) return 0;
    std::vector<CopyPiece> pc;
    for (auto& F : B->flows) pc.push_back(CopyPiece{nullptr, (unsigned int*)F.f, 4L * F.nb});
    for (auto& F : B->flows) pc.push_back(CopyPiece{nullptr, (unsigned int*)F.cg_q, 4L * F.nb});
    CKR(B->mem.alloc(&B->d_zero_pieces, pc.size()));
    return upload(B->d_zero_pieces, pc.data(), sizeof(CopyPiece) * pc.size());
}

// q = M x0 for r0 = b - q.  A zero start (warm start mode 0) needs no M
// application: q is cleared instead.
static int start_residual(ibpm_batch* B, int i) {
    hipStream_t s = B->ctx->c.stream;
    if (B->warm_mode == 0) {
        CKR(zero_pieces(B));
        launch_copy_pieces(B->d_zero_pieces + B->flows.size(), (int)B->flows.size(), s);
        return 0;
    }
    if (!B->d_start_pieces) {   // cg_d = f of every flow, one launch
        std::vector<CopyPiece> pc;
        for (auto& F : B->flows) pc.push_back(CopyPiece{(const unsigned int*)F.f, (unsigned int*)F.cg_d, 4L * F.nb});
        CKR(B->mem.alloc(&B->d_start_pieces, pc.size()));
        CKR(upload(B->d_start_pieces, pc.data(), sizeof(CopyPiece) * pc.size()));
    }
    launch_copy_pieces(B->d_start_pieces, (int)B->flows.size(), s);
    enqueue_M(B, i, B->d_smear_d[i], s);
    B->st_mapps += 1;
    return 0;
}

// Plain CG from x0 = f (ConjugateGradientSolver.cc:35-65): iterations in
// chunks, a device flag per flow; work for converged flows is skipped.
static int solve_cg(ibpm_batch* B, int i, int* iters) {
    hipStream_t s = B->ctx->c.stream;
    const int nfl = (int)B->flows.size();
    Probe* pr = B->probe;
    CKR(reset_states(B));
    CKR(start_residual(B, i));
    timed(pr, K_CG, 0, s, [&] { launch_cg(B->d_cg, nfl, 0, s); });
    std::vector<double> st;
    std::vector<char> active(nfl, 1);
    const int max_iter = 100000;
    const int chunk = pr ? 1 : B->cg_chunk;   // measuring: exact active flows per launch
    auto iterate = [&] {
        enqueue_M(B, i, B->d_smear_d[i], s, pr ? &active : nullptr);
        timed(pr, K_CG, 0, s, [&] { launch_cg(B->d_cg, nfl, 1, s); });
        B->st_mapps += 1;
    };
    // (not measuring: each poll reads the states while the first iteration of
    // the next chunk runs; past convergence it is a no-op)
    for (int it = 0; it < max_iter;) {
        CKR(poll_begin(B));
        if (!pr) iterate();
        CKR(poll_wait(B, st));
        bool all = true;
        for (int k = 0; k < nfl; ++k) {
            active[k] = st[8 * k + 3] == 0.0;
            if (active[k]) all = false;
        }
        if (all) break;
        for (int c = pr ? 0 : 1; c < chunk; ++c) iterate();
        it += chunk;
    }
    CKR(read_states(B, st));
    for (int k = 0; k < nfl; ++k) iters[k] += (int)st[8 * k + 2];
    return 0;
}

// Preconditioned CG from x0 = f on the same residual test.  *fallback is set
// when a flow broke down or hit the iteration limit; plain CG then finishes
// the solve from where PCG stopped.
static int solve_pcg(ibpm_batch* B, int i, int* iters, bool* fallback) {
    hipStream_t s = B->ctx->c.stream;
    const int nfl = (int)B->flows.size();
    Probe* pr = B->probe;
    const int par = B->pc_par[i];
    PcgOps* ops = B->d_pcg[i][par];
    CKR(reset_states(B));
    CKR(start_residual(B, i));
    std::vector<char> active(nfl, 1);
    auto precondition = [&](int begin) {
        double by = 0;   // P (n x n), its input and output, FP64, of the flows still iterating
        if (pr)
            for (int k = 0; k < nfl; ++k)
                if (active[k]) by += 8.0 * (4.0 * B->flows[k].nb * B->flows[k].nb + 4.0 * B->flows[k].nb);
        timed(pr, K_MINV, by, s, [&] { launch_minv(B->d_pminv[i], nfl, B->max_n, s); });
        timed(pr, K_CG, 0, s, [&] { launch_pcg_dir(ops, nfl, begin, s); });
    };
    timed(pr, K_CG, 0, s, [&] { launch_pcg_begin(ops, nfl, s); });
    precondition(1);
    // the first poll comes after as many iterations as the last solve of
    // this substep took, with nothing queued behind it: a solve that
    // converges there issues no iteration past convergence (each would be a
    // chain of no-op launches); later polls keep one iteration queued
    // (measuring: one iteration per poll, so the active flows of every
    // launch are known exactly)
    const int limit = std::max(100, 2 * B->max_n);
    int chunk = pr ? 0 : std::max(0, B->pc_last[i]), done = 0;
    bool first = chunk > 0;
    std::vector<double> st;
    auto iterate = [&] {
        enqueue_M(B, i, B->d_smear_d[i], s, pr ? &active : nullptr);
        timed(pr, K_CG, 0, s, [&] { launch_pcg_step(ops, nfl, s); });
        precondition(0);
        B->st_mapps += 1;
    };
    bool ahead = false;   // an iteration queued after the states being read
    for (;;) {
        for (int c = 0; c < chunk; ++c) iterate();
        done += chunk;
        if (pr) {   // measuring: the states of every launch exactly
            chunk = 1;
            CKR(read_states(B, st));
        } else if (first) {   // the expected count: read the states with nothing queued
            first = false;
            chunk = 0;
            CKR(read_states(B, st));
        } else {    // the next iteration runs while the host waits for the states
            CKR(poll_begin(B));
            iterate();
            ahead = true;
            done += 1;
            // from now on each round reads the states after the iteration
            // queued ahead in the round before and queues one more, so at most
            // one iteration past convergence is issued
            chunk = 0;
            CKR(poll_wait(B, st));
        }
        bool all = true;
        for (int k = 0; k < nfl; ++k) {
            active[k] = st[8 * k + 3] == 0.0 && st[8 * k + 4] == 0.0;
            if (active[k]) all = false;
        }
        if (all) break;   // the iteration ahead found every flow finished: a no-op
        if (done >= limit) {
            if (ahead) CKR(read_states(B, st));   // counts and flags after the iteration ahead
            break;
        }
    }
    B->pc_par[i] = 1 - par;
    *fallback = false;
    for (int k = 0; k < nfl; ++k) {
        iters[k] += (int)st[8 * k + 2];
        if (st[8 * k + 3] == 0.0) *fallback = true;
    }
    return 0;
}

// One substep with CG (moving bodies): host-side kinematics and regulariser,
// then CG iterations in chunks with a device-side convergence flag
// (ConjugateGradientSolver.cc:35-65; iterations after convergence are no-ops).
static int substep_cg(ibpm_batch* B, int i) {
    hipStream_t s = B->ctx->c.stream;
    const int nfl = (int)B->flows.size();
    const double t_kin = now_s();
    // updateOperators(t + c_n dt), IBSolver.cc:164-166: kinematics and
    // regulariser lists of the moving flows, computed during the previous
    // substep (kinematics_ahead) or here on host threads, then the uploads
    std::vector<int> moving;
    for (int k = 0; k < nfl; ++k)
        if (!B->flows[k].geom.stationary) moving.push_back(k);
    parallel_for((int)moving.size(), [&](int q) {
        Flow& F = B->flows[moving[q]];
        kinematics_at(F, B->time + B->cn[i] * F.p.dt);
    });
    // uploads of lists and descriptors queued on the stream behind the previous
    // substep's kernels, which read the old ones, while the GPU runs them
    {
        struct Guard {
            ibpm_batch* b;
            ~Guard() { b->async_up = false; }
        } guard{B};
        CKR(async_prep_begin(B));
        for (int k : moving) CKR(upload_regularizer(B, B->flows[k]));
        for (auto& F : B->flows)
            if (F.ubf) CKR(move_base_flow(B, F, B->time + B->cn[i] * F.p.dt));
        CKR(upload_regops(B));
        CKR(plan_local_helm(B, i));
        CKR(restrict_s2(B, i));
        CKR(async_prep_end(B));
    }
    B->st_host_kin_s += now_s() - t_kin;
    Probe* pr = B->probe;
    timed(pr, K_CLEAR, 0, s, [&] { launch_clear_nodes(B->d_clear, nfl, B->max_nodes, s); });
    enqueue_rhs(B, i, s);
    // P's refresh at this substep's points, beside the step solves (whose
    // transforms leave CUs idle; the stencils above fill the GPU)
    if (B->pc_mode && !B->pc_broken && B->pc_valid && B->pc_build_levels == 3) CKR(refresh_precond(B, i));
    run_solve(B->S1[i], s, pr);
    run_solve(B->S2[i], s, pr);
    timed(pr, K_INTERP, pr ? reg_bytes(B, false, true) : 0, s,
          [&] { launch_interp(B->d_interp_c, nfl, B->max_comp, s); });
    // the next substep's kinematics on host threads while the GPU runs this
    // substep's solves (its time exactly as that substep will form it)
    {
        const bool next_step = i + 1 == B->nsub;
        const int ni = next_step ? 0 : i + 1;
        const double t_next = next_step ? B->time + B->flows[0].p.dt : B->time;
        parallel_for((int)moving.size(), [&](int q) {
            Flow& F = B->flows[moving[q]];
            kinematics_ahead(F, t_next + B->cn[ni] * F.p.dt);
        });
    }
    // x0: previous f (ProjectionSolver.cc:72 warm start), zero, or the
    // learned prediction; then r0 = rhs - M x0
    if (B->warm_mode == 0) {
        CKR(zero_pieces(B));
        launch_copy_pieces(B->d_zero_pieces, nfl, s);
    } else if (B->warm_mode == 2 && B->nn_ready && nn_history_ready(B)) {
        CKR(nn_infer_table(B));
        const NnOps* d = B->d_nn_inf + ((size_t)(B->ksub % B->ring) * B->nsub + i) * nfl;
        int maxnb = 1;
        for (auto& F : B->flows) maxnb = std::max(maxnb, F.nb);
        launch_nn_features(d, nfl, s);
        launch_nn_infer(d, nfl, maxnb, s);
    }
    std::vector<int> its(nfl, 0);
    CKR(join_refresh(B));
    if (B->pc_mode && !B->pc_broken && !B->pc_valid) CKR(build_precond(B, i));
    if (B->pc_mode && !B->pc_broken) {
        bool fallback = false;
        CKR(solve_pcg(B, i, its.data(), &fallback));
        // rebuild P once the iterations it costs above its first solves
        // exceed what a rebuild costs: one M application per column, times
        // pc_build_cost (below 1 for the finest-level build)
        int itmax = 0;
        for (int k = 0; k < nfl; ++k) itmax = std::max(itmax, its[k]);
        B->pc_last[i] = itmax;
        // (a refreshed P does not go stale: only a jump of more than two
        // iterations over the first solve, a refresh losing track, counts)
        const int slack = B->pc_build_levels == 3 ? 2 : 0;
        if (B->pc_base[i] < 0) B->pc_base[i] = itmax;
        else B->pc_excess += std::max(0, itmax - B->pc_base[i] - slack);
        if (B->pc_excess > B->max_n * B->pc_build_cost) B->pc_valid = false;
        if (fallback) {
            B->st_fallbacks += 1;
            CKR(solve_cg(B, i, its.data()));
        }
    } else {
        CKR(solve_cg(B, i, its.data()));
    }
    for (int k = 0; k < nfl; ++k) {
        B->flows[k].cg_iters[i] = its[k];
        B->st_iters += its[k];
    }
    B->st_solves += nfl;
    // training samples (features of this substep, target = converged f) and
    // the force/constraint history ring
    if (B->collect && nn_history_ready(B) && nn_samples_room(B)) {
        CK(hipStreamSynchronize(s));
        CKR(upload_nn(B, i, true));
        launch_nn_features(B->d_nn, nfl, s);
    }
    if (!B->d_ring_pieces) {   // f and b into ring slot r of every flow: pieces [r 2 nflows, (r + 1) 2 nflows)
        std::vector<CopyPiece> pc;
        for (int r = 0; r < B->ring; ++r)
            for (auto& F : B->flows) {
                const size_t n = 2 * F.nb + 1;
                pc.push_back(CopyPiece{(const unsigned int*)F.f, (unsigned int*)(F.ring_f + r * n), 4L * F.nb});
                pc.push_back(CopyPiece{(const unsigned int*)F.b, (unsigned int*)(F.ring_b + r * n), 4L * F.nb});
            }
        CKR(B->mem.alloc(&B->d_ring_pieces, pc.size()));
        CKR(upload(B->d_ring_pieces, pc.data(), sizeof(CopyPiece) * pc.size()));
    }
    launch_copy_pieces(B->d_ring_pieces + (size_t)(B->ksub % B->ring) * 2 * nfl, 2 * nfl, s);
    B->ksub += 1;
    timed(pr, K_SMEAR, pr ? reg_bytes(B, true, false) : 0, s,
          [&] { launch_smear_curl(B->d_smear_f[i], nfl, B->max_nodes, s); });
    if (B->helm_on[i]) {
        // omega = omega* - A^-1 c with c = beta B f local like M's right-hand
        // side: the same sweeps (plan_local_helm), then the subtraction over
        // the boxes; off them omega already holds omega* (the S1 solve wrote it)
        for (int lev = 1; lev < B->S3[i].nlev; ++lev)
            timed(pr, K_HELM, 0, s, [&] { launch_helm_coarsify(B->d_helm3, B->helm_n, B->helm_cblocks, lev, s); });
        const int nl = helm_launches(B->helm_sweeps);
        double by = 0;
        for (double b : B->helm_bytes) by += b;
        launch_helm_sweeps(B->d_helm3, B->helm_n, B->helm_blocks, B->helm_sweeps, s,
                           [&](const std::function<void()>& fn) { timed(pr, K_HELM, by / nl, s, fn); });
        timed(pr, K_HELM, 0, s, [&] { launch_helm_epilogue(B->d_helm3, B->helm_n, B->helm_ablocks, s); });
    } else {
        run_solve(B->S3[i], s, pr);
    }
    enqueue_refresh(B, s);
    return 0;
}

// Cholesky factor M = L L^T from M's upper triangle exactly as
// CholeskySolver.cc:89-109 (descending-k sums): d = diagonal of L, lower =
// strict lower triangle, row-major (lower[r*n + c], c < r).  Mt is M
// column-major (Mt[c*n + r] = M(r, c)).  Returns -1, or the row where the
// matrix is not positive definite.
static int cholesky_factor(int n, const std::vector<double>& Mt, std::vector<double>& lower, std::vector<double>& d) {
    lower.assign((size_t)n * n, 0.0);
    d.assign(n, 0.0);
    auto M = [&](int r, int c) { return Mt[(size_t)c * n + r]; };
    for (int r = 0; r < n; ++r) {
        for (int c = r; c < n; ++c) {
            double sum = M(r, c);
            for (int k = r - 1; k >= 0; --k) sum -= lower[(size_t)r * n + k] * lower[(size_t)c * n + k];
            if (c == r) {
                if (!(sum > 0)) return r;
                d[r] = std::sqrt(sum);
            } else {
                lower[(size_t)c * n + r] = sum / d[r];
            }
        }
    }
    return -1;
}

// M^-1 row-major: CholeskySolver::Minv (:200-224) applied to every unit vector.
static void cholesky_assemble(int n, const std::vector<double>& lower, const std::vector<double>& d,
                              std::vector<double>& out) {
    out.assign((size_t)n * n, 0.0);
    std::vector<double> x(n);
    for (int col = 0; col < n; ++col) {
        for (int r = 0; r < n; ++r) {
            double sum = (r == col) ? 1.0 : 0.0;
            for (int k = r - 1; k >= 0; --k) sum -= lower[(size_t)r * n + k] * x[k];
            x[r] = sum / d[r];
        }
        for (int r = n - 1; r >= 0; --r) {
            double sum = x[r];
            for (int k = r + 1; k < n; ++k) sum -= lower[(size_t)k * n + r] * x[k];
            x[r] = sum / d[r];
        }
        for (int r = 0; r < n; ++r) out[(size_t)r * n + col] = x[r];
    }
}

// The preconditioner's P = M^-1, in place, from M's upper triangle as
// cholesky_factor reads it (Mt column-major in, P row-major out), in an order
// whose inner loops are contiguous row updates: U^T U = M by rows of U
// (right-looking), W = U^-1 by rows from the last, P = W W^T from W's rows and
// its transpose's.  About n^3/2 multiply-adds against the 4n^3/3 of the
// factor and unit-vector substitutions.  P only steers the iterations, so its
// rounding need not be the reference's (the Cholesky force solve keeps
// cholesky_factor / cholesky_assemble).  Returns -1, or the row where M is not
// positive definite.
static int spd_inverse_rows(int n, std::vector<double>& Mt) {
    std::vector<double> U((size_t)n * n, 0.0), W((size_t)n * n, 0.0);
    for (int r = 0; r < n; ++r)
        for (int c = r; c < n; ++c) U[(size_t)r * n + c] = Mt[(size_t)c * n + r];
    for (int k = 0; k < n; ++k) {
        double* uk = &U[(size_t)k * n];
        if (!(uk[k] > 0)) return k;
        const double d = std::sqrt(uk[k]), inv = 1.0 / d;
        uk[k] = d;
        for (int c = k + 1; c < n; ++c) uk[c] *= inv;
        for (int r = k + 1; r < n; ++r) {
            const double l = uk[r];
            double* ur = &U[(size_t)r * n];
            for (int c = r; c < n; ++c) ur[c] -= l * uk[c];
        }
    }
    // U W = I: row i of W is (e_i - sum over k > i of U[i][k] W[k]) / U[i][i]
    for (int i = n - 1; i >= 0; --i) {
        double* wi = &W[(size_t)i * n];
        const double* ui = &U[(size_t)i * n];
        wi[i] = 1.0;
        for (int k = i + 1; k < n; ++k) {
            const double u = ui[k];
            const double* wk = &W[(size_t)k * n];
            for (int j = k; j < n; ++j) wi[j] -= u * wk[j];
        }
        const double inv = 1.0 / ui[i];
        for (int j = i; j < n; ++j) wi[j] *= inv;
    }
    // P[i][j] = sum over k >= i of W[i][k] W[j][k] for j <= i (U reused as W^T), then mirrored
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) U[(size_t)k * n + j] = W[(size_t)j * n + k];
    std::fill(Mt.begin(), Mt.end(), 0.0);
    for (int i = 0; i < n; ++i) {
        double* pi = &Mt[(size_t)i * n];
        for (int k = i; k < n; ++k) {
            const double w = W[(size_t)i * n + k];
            const double* t = &U[(size_t)k * n];
            for (int j = 0; j <= i; ++j) pi[j] += w * t[j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) Mt[(size_t)j * n + i] = Mt[(size_t)i * n + j];
    return -1;
}

// In place: Mt (M column-major) -> M^-1 row-major.
static int cholesky_inverse(int n, std::vector<double>& Mt) {
    std::vector<double> lower, d;
    const int bad = cholesky_factor(n, Mt, lower, d);
    if (bad >= 0) return bad;
    cholesky_assemble(n, lower, d, Mt);
    return -1;
}

// The reference's factor cache <basename>_<substep, 2 digits>.cholesky
// (IBSolver.cc:77-108): size, alpha*beta, the diagonal, "#", then the strict
// lower triangle row by row, one value per line at 17 significant digits
// (CholeskySolver.cc:162-197).  Reading succeeds only for the same size and
// the same alpha*beta (CholeskySolver.cc:113-158).
static std::string cholesky_path(const std::string& base, int substep) {
    char num[16];
    std::snprintf(num, sizeof num, "_%02d", substep + 1);
    return base + num + ".cholesky";
}

static bool cholesky_write(const std::string& path, int n, double alpha_beta, const std::vector<double>& lower,
                           const std::vector<double>& d) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    std::fprintf(fp, "%d\n%.17g\n", n, alpha_beta);
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%.17g\n", d[i]);
    std::fprintf(fp, "#\n");
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) std::fprintf(fp, "%.17g\n", lower[(size_t)i * n + j]);
    return std::fclose(fp) == 0;
}

static bool cholesky_read(const std::string& path, int n, double alpha_beta, std::vector<double>& lower,
                          std::vector<double>& d) {
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return false;
    int nin = 0;
    double ab = 0;
    bool ok = std::fscanf(fp, "%d", &nin) == 1 && nin == n && std::fscanf(fp, "%lf", &ab) == 1 && ab == alpha_beta;
    if (ok) {
        d.assign(n, 0.0);
        lower.assign((size_t)n * n, 0.0);
        for (int i = 0; ok && i < n; ++i) ok = std::fscanf(fp, "%lf", &d[i]) == 1;
        char c = 0;
        ok = ok && std::fscanf(fp, " %c", &c) == 1 && c == '#';
        for (int i = 0; ok && i < n; ++i)
            for (int j = 0; ok && j < i; ++j) ok = std::fscanf(fp, "%lf", &lower[(size_t)i * n + j]) == 1;
    }
    std::fclose(fp);
    return ok;
}

// M of substep i for the current body positions, column by column on the GPU
// (CholeskySolver::init, CholeskySolver.cc:49-109), FP64 factorisation on the
// host and M^-1 = (L L^T)^-1 assembled with the reference's forward/back
// substitution (:200-224) applied to unit vectors.  The inverse goes to
// F.minv[i] (Cholesky force solve) or, with precond, to F.pinv.  Returns -5
// with precond when some M is not positive definite.
}  // namespace ibpmg
extern "C" {
static int solve_common(ibpm_ctx* c, int kind, int nx, int ny, int ngrid, double dx, double alpha,
                        const double* d_f, const double* d_bc, double* d_u, double in_scale, int precision);
}
namespace ibpmg {

static bool same_class(const ibpm_batch::KernTab& K, const Flow& F) {
    return K.nx == F.nx && K.ny == F.ny && K.ngrid == F.L && K.dx == F.dx && K.re == F.p.reynolds && K.dt == F.p.dt;
}

// Unit-flux nodes of a grid class's kernel table: a lattice over the level-0
// nodes its flows' points covered at set-up, nx/8 by ny/8 nodes apart (at
// least 8; coarser if more than 64 nodes), centred on that extent.  The
// level's boundary and the coarser levels make the operator depend on where a
// source sits, and a source within half a spacing of the points keeps that
// part of M (3e-3 of its spectrum with one source at the centre of the
// tandem plates' level, 1e-4 with a source every eighth of it).
static std::vector<int> table_sources(const std::vector<const Flow*>& fl) {
    const Flow& F0 = *fl[0];
    double imin = 1e300, imax = -1e300, jmin = 1e300, jmax = -1e300;
    for (const Flow* f : fl) {
        imin = std::min(imin, (f->bbox0[0] - f->p.xoffset) / f->dx);
        imax = std::max(imax, (f->bbox0[1] - f->p.xoffset) / f->dx);
        jmin = std::min(jmin, (f->bbox0[2] - f->p.yoffset) / f->dx);
        jmax = std::max(jmax, (f->bbox0[3] - f->p.yoffset) / f->dx);
    }
    const int i0 = (int)std::floor(imin), i1 = (int)std::ceil(imax), j0 = (int)std::floor(jmin), j1 = (int)std::ceil(jmax);
    int di = std::max(8, F0.nx / 8), dj = std::max(8, F0.ny / 8), ni = 1, nj = 1;
    for (;;) {
        ni = std::max(0, i1 - i0) / di + 1;
        nj = std::max(0, j1 - j0) / dj + 1;
        if (ni * nj <= 64) break;
        di *= 2;
        dj *= 2;
    }
    std::vector<int> src;
    for (int a = 0; a < ni; ++a)
        for (int b = 0; b < nj; ++b) {
            const int i = (i0 + i1) / 2 - ((ni - 1) * di) / 2 + a * di;
            const int j = (j0 + j1) / 2 - ((nj - 1) * dj) / 2 + b * dj;
            src.push_back(std::min(std::max(i, 1), F0.nx - 2));
            src.push_back(std::min(std::max(j, 1), F0.ny - 2));
        }
    return src;
}

// The kernel tables of every flow (once per grid class): for each substep i,
// G_i f = Curl L^-1 (-A_i^-1 Curl f) on level 0 (M_i = beta_i E G_i E^T,
// ProjectionSolver.cc:96-101 without beta) applied to a unit X flux and to a
// unit Y flux at each of the class's source nodes (table_sources), through
// the library's own curl, multi-domain Helmholtz solve and
// flux-from-vorticity; the level-0 part of each response is kept.
static int kernel_tables(ibpm_batch* B) {
    hipStream_t s = B->ctx->c.stream;
    for (auto& F : B->flows) {
        if (F.ktab[0]) continue;
        ibpm_batch::KernTab* T = nullptr;
        for (auto& K : B->ktabs)
            if (same_class(K, F)) T = &K;
        if (!T) {
            ibpm_batch::KernTab K{F.nx, F.ny, F.L, F.dx, F.p.reynolds, F.p.dt, {nullptr, nullptr, nullptr, nullptr},
                                  {}, nullptr};
            std::vector<const Flow*> members;
            for (const auto& G : B->flows)
                if (!G.geom.stationary && G.nx == F.nx && G.ny == F.ny && G.L == F.L && G.dx == F.dx &&
                    G.p.reynolds == F.p.reynolds && G.p.dt == F.p.dt)
                    members.push_back(&G);
            if (members.empty()) members.push_back(&F);
            K.src = table_sources(members);
            const int ns = (int)K.src.size() / 2;
            CKR(B->mem.alloc(&K.d_src, K.src.size()));
            CKR(upload(K.d_src, K.src.data(), sizeof(int) * K.src.size()));
            const long nf1 = 2L * F.nx * F.ny + F.nx + F.ny, nfx = (long)(F.nx + 1) * F.ny;
            DevMem tmp;
            double *q = nullptr, *w = nullptr, *u = nullptr, *r = nullptr;
            CKR(tmp.alloc(&q, nf1 * F.L)); CKR(tmp.alloc(&r, nf1 * F.L));
            CKR(tmp.alloc(&w, F.ns * F.L)); CKR(tmp.alloc(&u, F.ns * F.L));
            const double one = 1.0;
            for (int i = 0; i < B->nsub; ++i) {
                CKR(B->mem.alloc(&K.tab[i], 2 * nf1 * ns));
                for (int a = 0; a < ns; ++a) {
                    const int si = K.src[2 * a], sj = K.src[2 * a + 1];
                    const long unit[2] = {(long)si * F.ny + sj, nfx + (long)si * (F.ny + 1) + sj};
                    for (int comp = 0; comp < 2; ++comp) {
                        CK(hipMemsetAsync(q, 0, sizeof(double) * nf1 * F.L, s));
                        CK(hipMemcpyAsync(q + unit[comp], &one, sizeof(double), hipMemcpyHostToDevice, s));
                        CKR(ibpm_curl_flux(B->ctx, F.nx, F.ny, F.L, F.dx, q, w));
                        CKR(solve_common(B->ctx, 1, F.nx, F.ny, F.L, F.dx, -beta_of(B, F, i) / 2. / F.p.reynolds,
                                         w, nullptr, u, 1.0, 0));
                        CKR(ibpm_flux_from_vorticity(B->ctx, F.nx, F.ny, F.L, F.dx, 0.0, 0.0, u, r));
                        CK(hipMemcpyAsync(K.tab[i] + (2L * a + comp) * nf1, r, sizeof(double) * nf1,
                                          hipMemcpyDeviceToDevice, s));
                    }
                }
            }
            CK(hipStreamSynchronize(s));
            B->ktabs.push_back(K);
            T = &B->ktabs.back();
        }
        for (int i = 0; i < B->nsub; ++i) F.ktab[i] = T->tab[i];
        F.ksrc = T->d_src;
        F.nsrc = (int)T->src.size() / 2;
    }
    return 0;
}

// Assembly descriptors of the kernel-table matrix of substep i, scaled by
// scale(F), into each flow's mcol.
static std::vector<PcAsmOps> pc_asm_ops(ibpm_batch* B, int i, const std::function<double(const Flow&)>& scale) {
    std::vector<PcAsmOps> ops;
    for (auto& F : B->flows)
        ops.push_back(PcAsmOps{2 * F.nb, F.nx, F.ny, F.nsrc, F.ksrc, F.e_sep, F.ktab[i], scale(F), F.mcol});
    return ops;
}

static int build_inverse(ibpm_batch* B, int i, bool precond) {
    hipStream_t s = B->ctx->c.stream;
    const int nfl = (int)B->flows.size();
    int nmax = 0;
    for (auto& F : B->flows) nmax = std::max(nmax, 2 * F.nb);
    // factors of flows with a cache file that matches (Cholesky force solve only)
    std::vector<std::vector<double>> lower(nfl), diag(nfl);
    std::vector<char> cached(nfl, 0);
    bool all_cached = !precond;
    for (int k = 0; k < nfl; ++k) {
        Flow& F = B->flows[k];
        const double ab = (1. / F.p.reynolds) * beta_of(B, F, i);   // CholeskySolver.cc:38
        cached[k] = !precond && !F.chol_cache.empty() &&
                    cholesky_read(cholesky_path(F.chol_cache, i), 2 * F.nb, ab, lower[k], diag[k]);
        if (cached[k]) F.chol_loaded += 1;
        else all_cached = false;
    }
    // the preconditioner from M's finest-level form (pc_build_levels 1)
    const bool single = precond && B->pc_build_levels == 1 && B->helm_on[i];
    if (single && !B->m1l_built) {
        std::vector<SolveSpec> sp;
        for (auto& F : B->flows) {
            SolveSpec q = spec_of(F, 0, 0.0, F.m1, F.psi, -1.0, 0, nullptr);
            q.L = 1;            // level 0 alone: zero boundary values, no coarsify
            q.R = nullptr;
            q.skip = F.cg_state + 3;
            sp.push_back(q);
        }
        B->M1L.force_solve = true;
        CKR(build_solve_plan(&B->ctx->c, B->mem, sp, B->M1L));
        B->m1l_built = true;
    }
    if (single) {
        // the level-0 solve read as the force solve's: pass 1 in closed form
        // (zeros: no boundary values) off the local result's rows, passes 3 and
        // 2 over the node rows only
        for (auto& step : B->M1L.steps)
            for (auto& G : step) {
                bool changed = false;
                for (size_t j = 0; j < G.h.size(); ++j) {
                    EllJob& J = G.h[j];
                    const Flow& F = B->flows[G.jflow[j]];
                    const EllJob before = J;
                    whole_level(J);
                    if (!F.reg.node_idx.empty() && !F.hbox.empty()) {
                        int lo = INT_MAX, hi = INT_MIN;
                        for (int idx : F.reg.node_idx) {
                            lo = std::min(lo, idx / (F.ny - 1) + 1);
                            hi = std::max(hi, idx / (F.ny - 1) + 1);
                        }
                        J.sup_lo = F.hbox[0][0];
                        J.sup_hi = F.hbox[0][1];
                        J.sin1 = F.sin1;
                        read_rows_only(J, G, lo, hi, F.ringsin);
                    }
                    if (std::memcmp(&before, &J, sizeof(EllJob)) != 0) changed = true;
                }
                if (changed) {
                    CKR(upload(G.d, G.h.data(), sizeof(EllJob) * G.h.size()));
                    group_bytes(G);
                }
            }
    }
    const bool table = precond && B->pc_build_levels >= 2;
    if (table) {
        // the matrix from the kernel table at the current points (no M
        // applications): M_i = beta_i E G_i E^T
        CKR(kernel_tables(B));
        if (!B->d_pcasm_b) CKR(B->mem.alloc(&B->d_pcasm_b, (size_t)nfl));
        const auto ops = pc_asm_ops(B, i, [&](const Flow& F) { return beta_of(B, F, i); });
        CK(hipStreamSynchronize(s));
        CKR(upload(B->d_pcasm_b, ops.data(), sizeof(PcAsmOps) * nfl));
        launch_pc_assemble(B->d_pcasm_b, nfl, nmax, s);
    } else if (!all_cached) {
        // one smear descriptor per flow (the unit vector moves in place), one
        // interpolation descriptor per column and flow (column j of M into
        // mcol; column n is a dummy), so that the columns run without host
        // round trips
        std::vector<RegOps> sm(nfl), ip((size_t)(nmax + 1) * nfl);
        std::vector<UnitOps> uo(nfl);
        for (int k = 0; k < nfl; ++k) {
            Flow& F = B->flows[k];
            const int n = 2 * F.nb;
            sm[k] = make_regops(F, beta_of(B, F, i));
            sm[k].fin = F.unit; sm[k].cout = F.c;
            for (int j = 0; j <= nmax; ++j) {
                RegOps& r = ip[(size_t)j * nfl + k];
                r = make_regops(F, 0);
                r.psi = F.psi; r.b = nullptr;
                r.out = F.mcol + (size_t)std::min(j, n) * n;
            }
            uo[k] = UnitOps{F.unit, n};
            CK(hipMemsetAsync(F.unit, 0, sizeof(double) * (n + 1), s));
        }
        if (ip.size() > B->interp_cols_cap) {
            CKR(B->mem.alloc(&B->d_interp_cols, ip.size()));
            B->interp_cols_cap = ip.size();
        }
        if (!B->d_units) CKR(B->mem.alloc(&B->d_units, (size_t)nfl));
        CK(hipStreamSynchronize(s));
        CKR(upload(B->d_smear_d[i], sm.data(), sizeof(RegOps) * nfl));
        CKR(upload(B->d_interp_cols, ip.data(), sizeof(RegOps) * ip.size()));
        CKR(upload(B->d_units, uo.data(), sizeof(UnitOps) * nfl));
        for (int j = 0; j < nmax; ++j) {
            launch_unit_column(B->d_units, nfl, j, s);
            enqueue_M(B, i, B->d_smear_d[i], s, nullptr, B->d_interp_cols + (size_t)j * nfl, single);
        }
    }
    CK(hipStreamSynchronize(s));
    const double t_host = now_s();
    // factorisation and inverse assembly on the host, threads over flows
    std::vector<std::vector<double>> minv_h(nfl);
    std::vector<int> bad(nfl, -1);
    for (int k = 0; k < nfl; ++k) {
        Flow& F = B->flows[k];
        const int n = 2 * F.nb;
        minv_h[k].assign((size_t)n * n, 0.0);
        if (cached[k]) continue;
        CK(hipMemcpy(minv_h[k].data(), F.mcol, sizeof(double) * n * n, hipMemcpyDeviceToHost));
        CK(hipMemset(F.unit, 0, sizeof(double) * (n + 1)));
    }
    std::vector<std::thread> th;
    std::vector<char> saved(nfl, 1);
    const int hw = std::max(1u, std::thread::hardware_concurrency());
    const int nthreads = std::max(1, std::min({nfl, hw, 16}));
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t] {
            for (int k = t; k < nfl; k += nthreads) {
                Flow& F = B->flows[k];
                const int n = 2 * F.nb;
                if (precond) {   // P: the vectorised order
                    bad[k] = spd_inverse_rows(n, minv_h[k]);
                    continue;
                }
                if (!cached[k]) {
                    bad[k] = cholesky_factor(n, minv_h[k], lower[k], diag[k]);
                    if (bad[k] >= 0) continue;
                    if (!precond && !F.chol_cache.empty())   // IBSolver::save after init (ibpm.cc:291-296)
                        saved[k] = cholesky_write(cholesky_path(F.chol_cache, i), n,
                                                  (1. / F.p.reynolds) * beta_of(B, F, i), lower[k], diag[k]);
                }
                cholesky_assemble(n, lower[k], diag[k], minv_h[k]);
            }
        });
    for (auto& x : th) x.join();
    B->st_build_host_s += now_s() - t_host;
    for (int k = 0; k < nfl; ++k)
        if (!saved[k]) {
            set_error("could not write %s", cholesky_path(B->flows[k].chol_cache, i).c_str());
            return -1;
        }
    for (int k = 0; k < nfl; ++k) {
        if (bad[k] >= 0) {
            set_error("Cholesky: flow %d, matrix not positive definite at %d", k, bad[k]);
            CKR(upload_regops(B));
            return precond ? -5 : -4;
        }
        Flow& F = B->flows[k];
        const int n = 2 * F.nb;
        double*& dst = precond ? F.pinv : F.minv[i];
        if (!dst) CKR(B->mem.alloc(&dst, (size_t)n * n));
        CKR(upload(dst, minv_h[k].data(), sizeof(double) * n * n));
        if (precond) F.pbeta = beta_of(B, F, i);
    }
    if (!precond) {
        std::vector<MinvOps> mo(nfl);
        for (int k = 0; k < nfl; ++k) {
            Flow& F = B->flows[k];
            mo[k] = MinvOps{2 * F.nb, F.minv[i], F.rhs, F.f, 1.0, nullptr};
        }
        CKR(upload(B->d_minv[i], mo.data(), sizeof(MinvOps) * nfl));
    }
    return upload_regops(B);
}

static int build_cholesky(ibpm_batch* B) {
    for (int i = 0; i < B->nsub; ++i) CKR(build_inverse(B, i, false));
    return 0;
}

// The preconditioner for every substep from M of substep i at the current
// body positions: P_j = (beta_i / beta_j) M_i^-1, since M scales with the
// substep's beta.  An M that is not positive definite ends preconditioning
// for the batch (plain CG from then on).
static int build_precond(ibpm_batch* B, int i) {
    const int nfl = (int)B->flows.size();
    const double t0 = now_s();
    CKR(reset_states(B));   // the M applications below must not be skipped
    const int rc = build_inverse(B, i, true);
    B->pc_t_build = now_s() - t0;
    B->st_build_s += B->pc_t_build;
    B->st_builds += 1;
    if (B->pc_build_levels < 2) B->st_build_mapps += B->max_n;
    if (rc == -5) {
        B->pc_broken = true;
        return 0;
    }
    CKR(rc);
    for (int j = 0; j < B->nsub; ++j) {
        std::vector<MinvOps> mo(nfl);
        for (int k = 0; k < nfl; ++k) {
            Flow& F = B->flows[k];
            mo[k] = MinvOps{2 * F.nb, F.pinv, F.cg_t, F.cg_u, F.pbeta / beta_of(B, F, j), F.cg_state + 3};
        }
        CKR(upload(B->d_pminv[j], mo.data(), sizeof(MinvOps) * nfl));
    }
    if (B->pc_build_levels == 3) {
        // the refresh: mcol = pbeta E G_j E^T at substep j's points (P's
        // scale to substep j stays pbeta / beta_j), then per step pc_T = mcol
        // P, pc_X = 2 P - P pc_T, and the next step from pc_X back into P
        for (int j = 0; j < B->nsub; ++j) {
            if (!B->d_pcasm[j]) CKR(B->mem.alloc(&B->d_pcasm[j], (size_t)nfl));
            const auto ops = pc_asm_ops(B, j, [](const Flow& F) { return F.pbeta; });
            CKR(upload(B->d_pcasm[j], ops.data(), sizeof(PcAsmOps) * nfl));
        }
        std::vector<GemmOps> g1(2 * nfl), g2(2 * nfl);
        std::vector<CopyPiece> cp(nfl);
        for (int k = 0; k < nfl; ++k) {
            Flow& F = B->flows[k];
            const int n = 2 * F.nb;
            if (!F.pc_T) CKR(B->mem.alloc(&F.pc_T, (size_t)n * n));
            if (!F.pc_X) CKR(B->mem.alloc(&F.pc_X, (size_t)n * n));
            g1[k] = GemmOps{n, F.mcol, F.pinv, nullptr, F.pc_T, 1.0, 0.0};
            g2[k] = GemmOps{n, F.pinv, F.pc_T, F.pinv, F.pc_X, -1.0, 2.0};
            g1[nfl + k] = GemmOps{n, F.mcol, F.pc_X, nullptr, F.pc_T, 1.0, 0.0};
            g2[nfl + k] = GemmOps{n, F.pc_X, F.pc_T, F.pc_X, F.pinv, -1.0, 2.0};
            cp[k] = CopyPiece{(const unsigned int*)F.pc_X, (unsigned int*)F.pinv, 2L * n * n};
        }
        for (int h = 0; h < 2; ++h) {
            if (!B->d_ns1[h]) {
                CKR(B->mem.alloc(&B->d_ns1[h], (size_t)nfl));
                CKR(B->mem.alloc(&B->d_ns2[h], (size_t)nfl));
            }
            CKR(upload(B->d_ns1[h], g1.data() + h * nfl, sizeof(GemmOps) * nfl));
            CKR(upload(B->d_ns2[h], g2.data() + h * nfl, sizeof(GemmOps) * nfl));
        }
        if (!B->d_ns_copy) CKR(B->mem.alloc(&B->d_ns_copy, (size_t)nfl));
        CKR(upload(B->d_ns_copy, cp.data(), sizeof(CopyPiece) * nfl));
    }
    B->pc_valid = true;
    B->pc_excess = 0.0;
    for (int j = 0; j < 4; ++j) B->pc_base[j] = -1;
    return 0;
}

// pc_ns_steps Newton-Schulz steps of P toward the inverse of the
// kernel-table matrix at substep i's points, P <- P (2 I - M~ P)
// (symmetric), before the solve: each squares the error I - M~ P, so P follows
// the bodies from substep to substep at the cost of the assembly and two
// batched GEMMs per step.  The steps alternate between P and pc_X, so an even
// count copies nothing.
// Queued on the refresh stream behind the substep's uploads (the regulariser
// lists it reads), so that it overlaps the substep's right-hand side and step
// solves; join_refresh makes the stream wait for it before P is read.  P and
// the refresh's buffers are touched by nothing else in between: the previous
// solve's reads of P precede pc_ev_in, the next substep's uploads follow the
// join.
static int refresh_precond(ibpm_batch* B, int i) {
    hipStream_t s = B->ctx->c.stream;
    const int nfl = (int)B->flows.size();
    if (!B->pc_stream) {
        CK(hipStreamCreate(&B->pc_stream));
        CK(hipEventCreateWithFlags(&B->pc_ev_in, hipEventDisableTiming));
        CK(hipEventCreateWithFlags(&B->pc_ev_out, hipEventDisableTiming));
    }
    hipStream_t r = B->pc_stream;
    CK(hipEventRecord(B->pc_ev_in, s));
    CK(hipStreamWaitEvent(r, B->pc_ev_in, 0));
    Probe* pr = B->probe;
    timed(pr, K_PCASM, 0, r, [&] { launch_pc_assemble(B->d_pcasm[i], nfl, B->max_n, r); });
    for (int k = 0; k < B->pc_ns_steps; ++k)
        timed(pr, K_PCNS, 0, r, [&] {
            launch_dgemm(B->d_ns1[k & 1], nfl, B->max_n, false, r);
            launch_dgemm(B->d_ns2[k & 1], nfl, B->max_n, true, r);
            if ((k & 1) == 0 && k + 1 == B->pc_ns_steps) launch_copy_pieces(B->d_ns_copy, nfl, r);
        });
    CK(hipEventRecord(B->pc_ev_out, r));
    B->pc_pending = true;
    B->st_refreshes += 1;
    return 0;
}

static int join_refresh(ibpm_batch* B) {
    if (!B->pc_pending) return 0;
    CK(hipStreamWaitEvent(B->ctx->c.stream, B->pc_ev_out, 0));
    B->pc_pending = false;
    return 0;
}

// Buffers and descriptors of preconditioned CG (allocated on first use).
static int setup_pcg(ibpm_batch* B) {
    const int nfl = (int)B->flows.size();
    const int K = B->pc_pairs;
    for (auto& F : B->flows) {
        const int n = 2 * F.nb;
        if (!F.cg_t) {
            CKR(B->mem.alloc(&F.cg_t, n + 1)); CKR(B->mem.alloc(&F.cg_u, n + 1));
            CKR(B->mem.alloc(&F.lmp_g, kLmpMax * kLmpMax)); CKR(B->mem.alloc(&F.lmp_c, kLmpMax));
            CKR(B->mem.alloc(&F.lmp_cnt, 8));
        }
        CK(hipMemset(F.lmp_cnt, 0, sizeof(int) * 8));
        for (int i = 0; i < B->nsub; ++i)
            for (int p = 0; p < 2; ++p)
                if (K > 0 && !F.lmp_s[i][p]) {
                    CKR(B->mem.alloc(&F.lmp_s[i][p], (size_t)kLmpMax * n));
                    CKR(B->mem.alloc(&F.lmp_y[i][p], (size_t)kLmpMax * n));
                }
    }
    for (int i = 0; i < B->nsub; ++i) {
        if (!B->d_pminv[i]) CKR(B->mem.alloc(&B->d_pminv[i], nfl));
        for (int p = 0; p < 2; ++p) {
            std::vector<PcgOps> ops(nfl);
            for (int k = 0; k < nfl; ++k) {
                Flow& F = B->flows[k];
                PcgOps& o = ops[k];
                o.n = 2 * F.nb; o.kmax = K;
                o.x = F.f; o.r = F.cg_r; o.d = F.cg_d; o.q = F.cg_q; o.t = F.cg_t; o.u = F.cg_u;
                o.b = F.rhs; o.state = F.cg_state;
                o.rs = F.lmp_s[i][p]; o.ry = F.lmp_y[i][p]; o.rcnt = F.lmp_cnt + 2 * i + p;
                o.ws = F.lmp_s[i][1 - p]; o.wy = F.lmp_y[i][1 - p]; o.wcnt = F.lmp_cnt + 2 * i + 1 - p;
                o.g = F.lmp_g; o.c = F.lmp_c;
            }
            if (!B->d_pcg[i][p]) CKR(B->mem.alloc(&B->d_pcg[i][p], nfl));
            CKR(upload(B->d_pcg[i][p], ops.data(), sizeof(PcgOps) * nfl));
        }
        B->pc_par[i] = 0;
        B->pc_last[i] = 0;
    }
    B->pc_valid = false;
    B->pc_broken = false;
    return 0;
}

}  // namespace ibpmg

using namespace ibpmg;

// ===================================================================== C ABI
extern "C" {

const char* ibpm_last_error(void) { return g_err; }

int ibpm_geom_num_points(const char* geom_text) {
    Geometry g;
    std::string err;
    if (!g.parse(geom_text ? geom_text : "", err)) { set_error("%s", err.c_str()); return -1; }
    return g.npoints;
}

int ibpm_geom_info(const char* geom_text, int* npoints, int* nbodies, int* stationary, double* xc0, double* yc0) {
    Geometry g;
    std::string err;
    if (!g.parse(geom_text ? geom_text : "", err)) { set_error("%s", err.c_str()); return -1; }
    if (npoints) *npoints = g.npoints;
    if (nbodies) *nbodies = (int)g.bodies.size();
    if (stationary) *stationary = g.stationary ? 1 : 0;
    if (xc0) *xc0 = g.bodies.empty() ? 0.0 : g.bodies[0].xc;
    if (yc0) *yc0 = g.bodies.empty() ? 0.0 : g.bodies[0].yc;
    return 0;
}

int ibpm_geom_points(const char* geom_text, int ntimes, const double* times, double* xy, double* uv, long cap) {
    if (ntimes < 0 || (ntimes > 0 && (!times || !xy || !uv))) { set_error("ibpm_geom_points: bad arguments"); return -1; }
    Geometry g;
    std::string err;
    if (!g.parse(geom_text ? geom_text : "", err)) { set_error("%s", err.c_str()); return -1; }
    const long n = g.npoints;
    if (2 * n * ntimes > cap) {
        set_error("ibpm_geom_points: %ld values per array needed, room for %ld", 2 * n * ntimes, cap);
        return -1;
    }
    std::vector<double> x, y, u, v;
    for (int c = 0; c < ntimes; ++c) {
        g.move(times[c], x, y, u, v);
        double* pxy = xy + 2 * n * c;
        double* puv = uv + 2 * n * c;
        for (long k = 0; k < n; ++k) {
            pxy[k] = x[k]; pxy[n + k] = y[k];
            puv[k] = u[k]; puv[n + k] = v[k];
        }
    }
    return (int)n;
}

int ibpm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ibpm_ctx_create(int device, ibpm_ctx** out) {
    if (!out) { set_error("null output pointer"); return -1; }
    int n = ibpm_device_count();
    if (device < 0 || device >= n) { set_error("device %d not available (%d visible)", device, n); return -1; }
    auto* c = new ibpm_ctx();
    c->c.device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->c.stream) != hipSuccess) {
        set_error("could not create a stream on device %d", device);
        delete c;
        return -1;
    }
    if (scratch_attach(device) != 0) {   // the allocation-free entry point's array on this device
        (void)hipStreamDestroy(c->c.stream);
        delete c;
        return -1;
    }
    *out = c;
    return 0;
}

int ibpm_ctx_destroy(ibpm_ctx* c) {
    if (!c) return 0;
    (void)hipStreamSynchronize(c->c.stream);
    (void)hipStreamDestroy(c->c.stream);
    delete c;
    return 0;
}

int ibpm_ctx_synchronize(ibpm_ctx* c) {
    CK(hipStreamSynchronize(c->c.stream));
    return 0;
}

int ibpm_malloc(ibpm_ctx* c, size_t bytes, void** d) {
    (void)c;
    CK(hipMalloc(d, bytes ? bytes : 1));
    return 0;
}
int ibpm_free(ibpm_ctx* c, void* d) {
    (void)c;
    CK(hipFree(d));
    return 0;
}
int ibpm_memcpy_h2d(ibpm_ctx* c, void* d, const void* h, size_t bytes) {
    CK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->c.stream));
    CK(hipStreamSynchronize(c->c.stream));
    return 0;
}
int ibpm_memcpy_d2h(ibpm_ctx* c, void* h, const void* d, size_t bytes) {
    CK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->c.stream));
    CK(hipStreamSynchronize(c->c.stream));
    return 0;
}

static int check_grid(int nx, int ny, int ngrid) {
    if (nx < 2 || ny < 2 || ngrid < 1) { set_error("bad grid %dx%d, %d levels", nx, ny, ngrid); return -1; }
    if (ngrid > 1 && (nx % 4 || ny % 4)) {   // Grid.cc:72
        set_error("multi-domain grids need nx, ny divisible by 4 (got %dx%d)", nx, ny);
        return -1;
    }
    return 0;
}

int ibpm_dst2d(ibpm_ctx* c, int n1, int n2, const double* d_in, double* d_out) {
    // The 2-D DST is the solve pipeline with unit eigenvalues and no
    // normalisation: kind Poisson with a scaling table forced to 1 is not
    // available, so pass 1 and pass 3 are run back to back through T.
    CKR(check_grid(n1 + 1, n2 + 1, 1));
    DevMem mem;
    double *T = nullptr, *cosx = nullptr, *cosy = nullptr;
    CKR(mem.alloc(&T, (size_t)n1 * n2));
    CKR(cos_tables(mem, n1 + 1, &cosx));
    CKR(cos_tables(mem, n2 + 1, &cosy));
    EllJob J{};
    whole_level(J);
    J.nx = n1 + 1; J.ny = n2 + 1; J.lev = 0; J.ngrid = 1; J.in_scale = 1.0;
    J.f = d_in; J.T1 = T; J.T2 = T; J.u = d_out; J.norm = 1.0;
    J.ld1 = n1; J.ld2 = n2;   // unpadded: the second launch writes the caller's array
    CKR(get_plan(&c->c, n2, &J.pj));
    CKR(get_plan(&c->c, n1, &J.pi));
    // pass 1: rows along j -> T (transposed); a second pass-1 style launch over
    // the transposed rows transforms along i and transposes back.
    EllJob J2 = J;
    J2.nx = n2 + 1; J2.ny = n1 + 1; J2.f = T; J2.T1 = d_out; J2.pj = J.pi;
    J2.ld1 = n2; J2.ld2 = n1;
    EllJob* d = nullptr;
    CKR(mem.alloc(&d, 2));
    CK(hipMemcpyAsync(d, &J, sizeof(EllJob), hipMemcpyHostToDevice, c->c.stream));
    CK(hipMemcpyAsync(d + 1, &J2, sizeof(EllJob), hipMemcpyHostToDevice, c->c.stream));
    const bool r1 = pass_in_registers(J.pj.M, (n1 + 1) / 2), r2 = pass_in_registers(J2.pj.M, (n2 + 1) / 2);
    launch_ell_pass(1, d, 1, blocks_for_pass(J, 1, r1), lds_for_pass(J, 1, r1), c->c.stream, false, J.pj.M, r1);
    launch_ell_pass(1, d + 1, 1, blocks_for_pass(J2, 1, r2), lds_for_pass(J2, 1, r2), c->c.stream, false,
                    J2.pj.M, r2);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

static int solve_common(ibpm_ctx* c, int kind, int nx, int ny, int ngrid, double dx, double alpha,
                        const double* d_f, const double* d_bc, double* d_u, double in_scale = 1.0,
                        int precision = 0) {
    CKR(check_grid(nx, ny, ngrid));
    DevMem mem;
    const long ns = (long)(nx - 1) * (ny - 1);
    double *R = nullptr, *T1 = nullptr, *T2 = nullptr, *cosx = nullptr, *cosy = nullptr, *fin = nullptr;
    CKR(mem.alloc(&R, ns * ngrid));
    CKR(mem.alloc(&T1, (size_t)(ny - 1) * t1_ld(nx - 1)));
    CKR(mem.alloc(&T2, (size_t)(nx - 1) * t1_ld(ny - 1)));
    CKR(mem.alloc(&fin, ns * ngrid));
    CK(hipMemcpyAsync(fin, d_f, sizeof(double) * ns * ngrid, hipMemcpyDeviceToDevice, c->c.stream));
    CKR(cos_tables(mem, nx, &cosx));
    CKR(cos_tables(mem, ny, &cosy));
    SolveSpec s{};
    s.nx = nx; s.ny = ny; s.L = ngrid; s.dx = dx; s.kind = kind; s.alpha = alpha;
    s.f = fin; s.u = d_u; s.R = R; s.T1 = T1; s.T2 = T2; s.in_scale = in_scale; s.epilogue = 0;
    s.cosx = cosx; s.cosy = cosy; s.bcring = d_bc;
    SolvePlan P;
    CKR(build_solve_plan(&c->c, mem, {s}, P));
    if (precision >= 1) set_plan_fp32(P, true);
    if (precision == 2) {
        CKR(tune_plan_gemm(&c->c, mem, P, 4));
        for (auto& step : P.steps)
            for (auto& G : step)
                if (!G.gemm) {
                    set_error("GEMM-DST: square levels of at most 511 unknowns per side only (got %dx%d)", nx, ny);
                    return -1;
                }
    }
    run_solve(P, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_elliptic2d_solve(ibpm_ctx* c, int kind, int nx, int ny, double dx, double alpha,
                          const double* d_f, const double* d_bc, double* d_u) {
    return solve_common(c, kind, nx, ny, 1, dx, alpha, d_f, d_bc, d_u);
}

int ibpm_elliptic_solve(ibpm_ctx* c, int kind, int nx, int ny, int ngrid, double dx, double alpha,
                        const double* d_f, double* d_u) {
    return solve_common(c, kind, nx, ny, ngrid, dx, alpha, d_f, nullptr, d_u);
}

int ibpm_elliptic_solve_precision(ibpm_ctx* c, int kind, int nx, int ny, int ngrid, double dx, double alpha,
                                  const double* d_f, double* d_u, int precision) {
    if (precision < 0 || precision > 2) {
        set_error("precision %d: 0 (FP64 transforms), 1 (FP32 transforms) or 2 (split-BF16 GEMM-DST)", precision);
        return -1;
    }
    return solve_common(c, kind, nx, ny, ngrid, dx, alpha, d_f, nullptr, d_u, 1.0, precision);
}

// ibpm_elliptic_solve_noalloc: the multi-domain solve of ibpm_elliptic_solve
// with every buffer of its own (right-hand-side copy, coarsify buffer,
// transpose buffers, cosine and twiddle tables, job descriptors) carved from
// the device's fixed scratch array instead of allocated.  The host values of
// the tables and descriptors are staged in static memory under a lock, which
// is held until the stream has drained.
static constexpr int kNoAllocMaxLevels = 16;
struct NoAllocStage {
    double cosx[kFftLds / 2], cosy[kFftLds / 2];
    double2 twj[kFftLds], t16j[256], twi[kFftLds], t16i[256];
    EllJob ej[kNoAllocMaxLevels];
    CoarsifyJob cj[kNoAllocMaxLevels];
};
static std::mutex g_noalloc_mu;
static NoAllocStage g_noalloc_stage;

int ibpm_elliptic_solve_noalloc(ibpm_ctx* c, int kind, int nx, int ny, int ngrid, double dx, double alpha,
                                const double* d_f, double* d_u) {
    if (!c || !d_f || !d_u) { set_error("ibpm_elliptic_solve_noalloc: null context or array"); return -1; }
    if (kind != 0 && kind != 1) { set_error("ibpm_elliptic_solve_noalloc: kind %d is not 0 or 1", kind); return -1; }
    CKR(check_grid(nx, ny, ngrid));
    if (ngrid > kNoAllocMaxLevels) {
        set_error("ibpm_elliptic_solve_noalloc: %d levels exceed the limit of %d", ngrid, kNoAllocMaxLevels);
        return -2;
    }
    FftPlan pj{}, pi{};
    CKR(plan_shape(ny - 1, &pj));
    CKR(plan_shape(nx - 1, &pi));
    const int L = ngrid;
    const long ns = (long)(nx - 1) * (ny - 1);
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_fin = take(8 * ns * L), o_R = take(8 * ns * L);
    const size_t o_T1 = take(8 * (size_t)(ny - 1) * t1_ld(nx - 1)), o_T2 = take(8 * (size_t)(nx - 1) * t1_ld(ny - 1));
    const size_t zero_lo = o_R, zero_hi = off;   // R, T1 and T2 start at zero, as freshly allocated ones
    const size_t o_cx = take(8 * (size_t)(nx - 1)), o_cy = take(8 * (size_t)(ny - 1));
    const size_t o_twj = take(16 * (size_t)pj.M), o_t16j = take(16 * 256);
    const size_t o_twi = take(16 * (size_t)pi.M), o_t16i = take(16 * 256);
    const size_t o_ej = take(sizeof(EllJob) * L), o_cj = take(sizeof(CoarsifyJob) * L);
    if (off > kScratchBytes) {
        set_error("ibpm_elliptic_solve_noalloc: a %dx%d grid of %d levels needs %zu bytes of scratch; "
                  "the fixed scratch holds %zu", nx, ny, ngrid, off, kScratchBytes);
        return -2;
    }
    void* base = nullptr;
    CKR(scratch_device_ptr(c->c.device, &base));
    char* b = (char*)base;
    double* fin = (double*)(b + o_fin);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoAllocStage& H = g_noalloc_stage;
    cos_fill(nx, H.cosx);
    cos_fill(ny, H.cosy);
    plan_tables(pj.M, H.twj, H.t16j);
    plan_tables(pi.M, H.twi, H.t16i);
    pj.tw = (const double2*)(b + o_twj);
    pj.tw16 = plan_has_tw16(pj.M) ? (const double2*)(b + o_t16j) : nullptr;
    pi.tw = (const double2*)(b + o_twi);
    pi.tw16 = plan_has_tw16(pi.M) ? (const double2*)(b + o_t16i) : nullptr;
    SolveSpec sp{};
    sp.nx = nx; sp.ny = ny; sp.L = L; sp.dx = dx; sp.kind = kind; sp.alpha = alpha;
    sp.f = fin; sp.u = d_u; sp.R = (double*)(b + o_R); sp.T1 = (double*)(b + o_T1); sp.T2 = (double*)(b + o_T2);
    sp.in_scale = 1.0; sp.epilogue = 0;
    sp.cosx = (const double*)(b + o_cx); sp.cosy = (const double*)(b + o_cy);
    for (int st = 0; st < L; ++st) {   // coarsest level first
        H.ej[st] = EllJob{};
        fill_ell_job(sp, L - 1 - st, H.ej[st]);
        H.ej[st].pj = pj;
        H.ej[st].pi = pi;
    }
    for (int lev = 1; lev < L; ++lev) H.cj[lev - 1] = CoarsifyJob{nx, ny, lev, sp.f, sp.R, nullptr};
    hipStream_t s = c->c.stream;
    EllJob* d_ej = (EllJob*)(b + o_ej);
    CoarsifyJob* d_cj = (CoarsifyJob*)(b + o_cj);
    CK(hipMemsetAsync(b + zero_lo, 0, zero_hi - zero_lo, s));
    CK(hipMemcpyAsync(fin, d_f, 8 * ns * L, hipMemcpyDeviceToDevice, s));
    CK(hipMemcpyAsync(b + o_cx, H.cosx, 8 * (size_t)(nx - 1), hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(b + o_cy, H.cosy, 8 * (size_t)(ny - 1), hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(b + o_twj, H.twj, 16 * (size_t)pj.M, hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(b + o_t16j, H.t16j, 16 * 256, hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(b + o_twi, H.twi, 16 * (size_t)pi.M, hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(b + o_t16i, H.t16i, 16 * 256, hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(d_ej, H.ej, sizeof(EllJob) * L, hipMemcpyHostToDevice, s));
    if (L > 1) CK(hipMemcpyAsync(d_cj, H.cj, sizeof(CoarsifyJob) * (L - 1), hipMemcpyHostToDevice, s));
    for (int lev = 1; lev < L; ++lev) launch_coarsify(d_cj + (lev - 1), 1, nx / 2 - 1, s, false);
    for (int st = 0; st < L; ++st) {
        bool regs[3];
        int maxb[3], lds[3];
        pass_shape(H.ej[st], regs, maxb, lds);
        for (int pass = 1; pass <= 3; ++pass)
            launch_ell_pass(pass, d_ej + st, 1, maxb[pass - 1], lds[pass - 1], s, false,
                            pass == 2 ? pi.M : pj.M, regs[pass - 1]);
    }
    CK(hipGetLastError());
    CK(hipStreamSynchronize(s));
    return 0;
}

// Stand-alone stencil entry points run the batch kernels on a one-flow
// descriptor with scratch buffers.
static int ops_common(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, FlowOps& o, DevMem& mem) {
    CKR(check_grid(nx, ny, ngrid));
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    const long nf = (2L * nx * ny + nx + ny) * ngrid;
    o.nx = nx; o.ny = ny; o.L = ngrid; o.dx = dx;
    if (!o.U) CKR(mem.alloc(&o.U, ns));
    if (!o.V) CKR(mem.alloc(&o.V, ns));
    if (!o.cross) CKR(mem.alloc(&o.cross, nf));
    if (!o.a) CKR(mem.alloc(&o.a, ns));
    if (!o.nprev) CKR(mem.alloc(&o.nprev, ns));
    (void)c;
    return 0;
}

int ibpm_curl_flux(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, const double* d_q, double* d_w) {
    // N = Curl(cross) of the rhs kernel with omega = 0: a = dt * an * N = N.
    DevMem mem;
    FlowOps o{};
    double* zero = nullptr;
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    CKR(mem.alloc(&zero, ns));
    o.cross = const_cast<double*>(d_q);
    o.omega = zero;
    o.a = d_w;
    o.lap_coef = 0; o.an = 1; o.bn = 0; o.dt = 1; o.w_coef = 0;
    CKR(ops_common(c, nx, ny, ngrid, dx, o, mem));
    FlowOps* d = nullptr;
    CKR(mem.alloc(&d, 1));
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    launch_rhs(d, 1, ns, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_laplacian(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, const double* d_w, double* d_out) {
    // rhs kernel with a = Lap(omega): lap_coef 1, no nonlinear term, no omega.
    DevMem mem;
    FlowOps o{};
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    const long nf = (2L * nx * ny + nx + ny) * ngrid;
    double* zc = nullptr;
    CKR(mem.alloc(&zc, nf));
    o.cross = zc;
    o.omega = const_cast<double*>(d_w);
    o.a = d_out;
    o.lap_coef = 1; o.an = 0; o.bn = 0; o.dt = 1; o.w_coef = 0;
    CKR(ops_common(c, nx, ny, ngrid, dx, o, mem));
    FlowOps* d = nullptr;
    CKR(mem.alloc(&d, 1));
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    launch_rhs(d, 1, ns, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_curl_scalar(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, double u0, double v0,
                     const double* d_w, double* d_q) {
    DevMem mem;
    FlowOps o{};
    o.psi = const_cast<double*>(d_w);
    o.q = d_q;
    o.ubase = u0; o.vbase = v0;
    CKR(ops_common(c, nx, ny, ngrid, dx, o, mem));
    FlowOps* d = nullptr;
    CKR(mem.alloc(&d, 1));
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    launch_curl_q(d, 1, (long)(nx + 1) * (ny + 1) * ngrid, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_nonlinear(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, const double* d_q,
                   const double* d_w, double* d_n) {
    DevMem mem;
    FlowOps o{};
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    const long nf1 = 2L * nx * ny + nx + ny;
    double* zero = nullptr;
    CKR(mem.alloc(&zero, ns));
    o.q = const_cast<double*>(d_q);
    o.omega = const_cast<double*>(d_w);
    CKR(ops_common(c, nx, ny, ngrid, dx, o, mem));
    FlowOps* d = nullptr;
    CKR(mem.alloc(&d, 2));
    FlowOps o2 = o;   // curl of cross: rhs kernel with omega = 0, dt = an = 1
    o2.omega = zero; o2.a = d_n; o2.lap_coef = 0; o2.an = 1; o2.bn = 0; o2.dt = 1; o2.w_coef = 0;
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    CK(hipMemcpyAsync(d + 1, &o2, sizeof(o2), hipMemcpyHostToDevice, c->c.stream));
    launch_vel(d, 1, ns, c->c.stream);
    for (int lev = 0; lev < ngrid; ++lev) launch_cross(d, 1, nf1, lev, c->c.stream);
    launch_rhs(d + 1, 1, ns, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_cross_product(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, const double* d_q, const double* d_w,
                       double* d_out) {
    DevMem mem;
    FlowOps o{};
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    const long nf1 = 2L * nx * ny + nx + ny;
    o.q = const_cast<double*>(d_q);
    o.omega = const_cast<double*>(d_w);
    o.cross = d_out;
    CKR(ops_common(c, nx, ny, ngrid, dx, o, mem));
    FlowOps* d = nullptr;
    CKR(mem.alloc(&d, 1));
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    launch_vel(d, 1, ns, c->c.stream);
    for (int lev = 0; lev < ngrid; ++lev) launch_cross(d, 1, nf1, lev, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

// Scalar::coarsify in place over levels 1..ngrid-1, in order (Scalar.cc:46-63):
// each level reads the finer one, already coarsified.
static void coarsify_in_place(int nx, int ngrid, const CoarsifyJob* d_jobs, hipStream_t s) {
    for (int lev = 1; lev < ngrid; ++lev) launch_coarsify(d_jobs + lev, 1, nx / 2 - 1, s);
}

int ibpm_cross_flux_flux(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, const double* d_q1,
                         const double* d_q2, double* d_out) {
    CKR(check_grid(nx, ny, ngrid));
    DevMem mem;
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    CrossFFOps o{nx, ny, ngrid, dx, d_q1, d_q2, d_out};
    std::vector<CoarsifyJob> jobs(ngrid);
    for (int lev = 0; lev < ngrid; ++lev) jobs[lev] = CoarsifyJob{nx, ny, lev, d_out, d_out, nullptr};
    CrossFFOps* d = nullptr;
    CoarsifyJob* dj = nullptr;
    CKR(mem.alloc(&d, 1));
    CKR(mem.alloc(&dj, ngrid));
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    CK(hipMemcpyAsync(dj, jobs.data(), sizeof(CoarsifyJob) * ngrid, hipMemcpyHostToDevice, c->c.stream));
    launch_cross_ff(d, 1, ns, c->c.stream);
    coarsify_in_place(nx, ngrid, dj, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_rhs(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, double dt, double nu, double an, double bn,
             const double* d_q, const double* d_w, const double* d_nprev, double* d_a, double* d_n) {
    DevMem mem;
    FlowOps o{};
    const long ns = (long)(nx - 1) * (ny - 1) * ngrid;
    const long nf1 = 2L * nx * ny + nx + ny;
    o.q = const_cast<double*>(d_q);
    o.omega = const_cast<double*>(d_w);
    o.a = d_a;
    o.nprev = d_n;   // holds N_prev on entry (when given) and N on exit
    o.lap_coef = 0.5 * nu * (an + bn); o.an = an; o.bn = bn; o.dt = dt; o.w_coef = 1.0;
    o.use_nprev = bn != 0;
    o.nprev_valid = d_nprev != nullptr;   // no N_prev yet: the reference uses N (IBSolver.cc:173-177)
    CKR(ops_common(c, nx, ny, ngrid, dx, o, mem));
    if (d_nprev && d_nprev != d_n)
        CK(hipMemcpyAsync(d_n, d_nprev, sizeof(double) * ns, hipMemcpyDeviceToDevice, c->c.stream));
    FlowOps* d = nullptr;
    CKR(mem.alloc(&d, 1));
    CK(hipMemcpyAsync(d, &o, sizeof(o), hipMemcpyHostToDevice, c->c.stream));
    launch_vel(d, 1, ns, c->c.stream);
    for (int lev = 0; lev < ngrid; ++lev) launch_cross(d, 1, nf1, lev, c->c.stream);
    launch_rhs(d, 1, ns, c->c.stream);
    CK(hipStreamSynchronize(c->c.stream));
    CK(hipGetLastError());
    return 0;
}

int ibpm_flux_from_vorticity(ibpm_ctx* c, int nx, int ny, int ngrid, double dx, double u0, double v0,
                             const double* d_w, double* d_q) {
    CKR(check_grid(nx, ny, ngrid));
    DevMem mem;
    double* psi = nullptr;
    CKR(mem.alloc(&psi, (size_t)(nx - 1) * (ny - 1) * ngrid));
    CKR(solve_common(c, 0, nx, ny, ngrid, dx, 0.0, d_w, nullptr, psi, -1.0));   // L psi = -coarsify(omega)
    return ibpm_curl_scalar(c, nx, ny, ngrid, dx, u0, v0, psi, d_q);
}

}  // extern "C"

// Standalone regulariser: association lists of one set of points.
struct ibpm_reg {
    ibpm_ctx* ctx = nullptr;
    int nx = 0, ny = 0, ngrid = 0, nb = 0, cap = 0;
    double dx = 0, x0 = 0, y0 = 0;
    ibpmg::DevMem mem;
    int *e_ptr = nullptr, *e_f = nullptr, *t_ptr = nullptr, *t_k = nullptr, *fos = nullptr;
    double *e_w = nullptr, *t_w = nullptr;
    ibpmg::RegFluxOps* d_ops = nullptr;
    int nslots = 0;
};

extern "C" {

int ibpm_reg_update(ibpm_reg* r, const double* h_xy) {
    if (!r || !h_xy) { set_error("null argument"); return -1; }
    std::vector<double> px(h_xy, h_xy + r->nb), py(h_xy + r->nb, h_xy + 2 * r->nb);
    RegLists L;
    build_reg_lists(r->nx, r->ny, r->dx, r->x0, r->y0, px, py, L);
    if ((int)L.e_f.size() > r->cap || (int)L.t_k.size() > r->cap) {
        set_error("regulariser list exceeds capacity");
        return -3;
    }
    CKR(upload(r->e_ptr, L.e_ptr.data(), sizeof(int) * L.e_ptr.size()));
    CKR(upload(r->e_f, L.e_f.data(), sizeof(int) * L.e_f.size()));
    CKR(upload(r->e_w, L.e_w.data(), sizeof(double) * L.e_w.size()));
    CKR(upload(r->t_ptr, L.t_ptr.data(), sizeof(int) * L.t_ptr.size()));
    CKR(upload(r->t_k, L.t_k.data(), sizeof(int) * L.t_k.size()));
    CKR(upload(r->t_w, L.t_w.data(), sizeof(double) * L.t_w.size()));
    CKR(upload(r->fos, L.flux_of_slot.data(), sizeof(int) * L.flux_of_slot.size()));
    r->nslots = (int)L.flux_of_slot.size();
    RegFluxOps o{r->nb, r->nslots, r->dx, r->e_ptr, r->e_f, r->e_w, r->t_ptr, r->t_k, r->t_w, r->fos};
    CKR(upload(r->d_ops, &o, sizeof(o)));
    return 0;
}

int ibpm_reg_create(ibpm_ctx* c, int nx, int ny, int ngrid, double length, double xoffset, double yoffset,
                    int npts, const double* h_xy, ibpm_reg** out) {
    if (!c || !out || npts < 1 || !h_xy) { set_error("bad regulariser arguments"); return -1; }
    CKR(check_grid(nx, ny, ngrid));
    auto r = std::make_unique<ibpm_reg>();
    r->ctx = c; r->nx = nx; r->ny = ny; r->ngrid = ngrid; r->nb = npts;
    r->dx = length / nx; r->x0 = xoffset; r->y0 = yoffset;
    r->cap = 18 * npts + 1;
    const int n = 2 * npts;
    CKR(r->mem.alloc(&r->e_ptr, n + 1)); CKR(r->mem.alloc(&r->e_f, r->cap)); CKR(r->mem.alloc(&r->e_w, r->cap));
    CKR(r->mem.alloc(&r->t_ptr, r->cap + 1)); CKR(r->mem.alloc(&r->t_k, r->cap));
    CKR(r->mem.alloc(&r->t_w, r->cap)); CKR(r->mem.alloc(&r->fos, r->cap));
    CKR(r->mem.alloc(&r->d_ops, 1));
    CKR(ibpm_reg_update(r.get(), h_xy));
    *out = r.release();
    return 0;
}

int ibpm_reg_destroy(ibpm_reg* r) {
    delete r;
    return 0;
}

int ibpm_reg_to_flux(ibpm_reg* r, const double* d_f, double* d_q) {
    if (!r) { set_error("null regulariser"); return -1; }
    hipStream_t s = r->ctx->c.stream;
    const size_t nf = (size_t)(2L * r->nx * r->ny + r->nx + r->ny) * r->ngrid;
    CK(hipMemsetAsync(d_q, 0, sizeof(double) * nf, s));   // coarse levels and untouched fluxes are 0
    launch_reg_to_flux(r->d_ops, r->nslots, d_f, d_q, s);
    CK(hipStreamSynchronize(s));
    CK(hipGetLastError());
    return 0;
}

int ibpm_reg_to_boundary(ibpm_reg* r, const double* d_q, double* d_f) {
    if (!r) { set_error("null regulariser"); return -1; }
    hipStream_t s = r->ctx->c.stream;
    launch_reg_to_boundary(r->d_ops, r->nb, d_q, d_f, s);
    CK(hipStreamSynchronize(s));
    CK(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ model
// One flow's NavierStokesModel operators over the stand-alone entry points:
// the regulariser of the geometry's points at the model's time, and the
// elliptic solves, curls and stencils of one grid.  The batch is the fast
// path; this object is the reference's operator-level interface.
}  // extern "C"

struct ibpm_model {
    ibpm_ctx* ctx = nullptr;
    ibpm_flow_params p{};
    std::string geom;
    int nb = 0;
    double dx = 0, u0 = 0, v0 = 0, time = 0;
    std::vector<double> xy, uv;     // points and velocities at `time`
    ibpm_reg* reg = nullptr;
    ibpmg::DevMem mem;
    double *d_q = nullptr, *d_w = nullptr, *d_u = nullptr, *d_f = nullptr, *d_y = nullptr;
    long ns = 0, nf = 0;
    ~ibpm_model() {
        if (reg) ibpm_reg_destroy(reg);
    }
};

namespace ibpmg {
static int model_points(ibpm_model* m, double t) {
    if (m->nb == 0) return 0;
    if (ibpm_geom_points(m->geom.c_str(), 1, &t, m->xy.data(), m->uv.data(), 2L * m->nb) != m->nb) {
        set_error("the geometry does not parse");
        return -1;
    }
    return 0;
}
}  // namespace ibpmg

extern "C" {

int ibpm_model_create(ibpm_ctx* c, const ibpm_flow_params* p, double t0, ibpm_model** out) {
    if (!c || !p || !out || !p->geom) { set_error("bad model arguments"); return -1; }
    CKR(check_grid(p->nx, p->ny, p->ngrid));
    if (!(p->reynolds > 0)) { set_error("Reynolds number must be positive"); return -1; }
    auto m = std::make_unique<ibpm_model>();
    m->ctx = c;
    m->p = *p;
    m->geom = p->geom;
    m->p.geom = nullptr;
    m->nb = ibpm_geom_num_points(m->geom.c_str());
    if (m->nb < 0) { set_error("the geometry does not parse"); return -1; }
    m->dx = p->length / p->nx;
    double sa, ca;   // BaseFlow(grid, 1, alpha): UniformFlow of magnitude 1 (Flux.cc:79-99)
    const double pi = 4. * std::atan(1.);
    sin_cos(p->alpha_deg * pi / 180., sa, ca);
    m->u0 = ca;
    m->v0 = sa;
    m->time = t0;
    m->ns = (long)(p->nx - 1) * (p->ny - 1) * p->ngrid;
    m->nf = (2L * p->nx * p->ny + p->nx + p->ny) * p->ngrid;
    CKR(m->mem.alloc(&m->d_q, m->nf));
    CKR(m->mem.alloc(&m->d_w, m->ns));
    CKR(m->mem.alloc(&m->d_u, m->ns));
    CKR(m->mem.alloc(&m->d_f, std::max(1, 2 * m->nb)));
    CKR(m->mem.alloc(&m->d_y, std::max(1, 2 * m->nb)));
    m->xy.assign(2 * (size_t)m->nb, 0.0);
    m->uv.assign(2 * (size_t)m->nb, 0.0);
    CKR(model_points(m.get(), t0));
    if (m->nb > 0)
        CKR(ibpm_reg_create(c, p->nx, p->ny, p->ngrid, p->length, p->xoffset, p->yoffset, m->nb, m->xy.data(),
                            &m->reg));
    *out = m.release();
    return 0;
}

int ibpm_model_destroy(ibpm_model* m) {
    delete m;
    return 0;
}

int ibpm_model_num_points(ibpm_model* m) { return m ? m->nb : -1; }

int ibpm_model_set_base_flow(ibpm_model* m, double mag, double alpha) {
    if (!m) { set_error("null model"); return -1; }
    double sa, ca;   // UniformFlow, Flux.cc:79-99
    sin_cos(alpha, sa, ca);
    m->u0 = mag * ca;
    m->v0 = mag * sa;
    return 0;
}

int ibpm_model_update_operators(ibpm_model* m, double time) {
    if (!m) { set_error("null model"); return -1; }
    m->time = time;
    CKR(model_points(m, time));
    if (m->reg) CKR(ibpm_reg_update(m->reg, m->xy.data()));
    return 0;
}

int ibpm_model_B(ibpm_model* m, const double* d_f, double* d_omega) {
    if (!m) { set_error("null model"); return -1; }
    const ibpm_flow_params& p = m->p;
    if (!m->reg) {   // no points: E^T f = 0
        CK(hipMemsetAsync(d_omega, 0, sizeof(double) * m->ns, m->ctx->c.stream));
        CK(hipStreamSynchronize(m->ctx->c.stream));
        return 0;
    }
    CKR(ibpm_reg_to_flux(m->reg, d_f, m->d_q));
    return ibpm_curl_flux(m->ctx, p.nx, p.ny, p.ngrid, m->dx, m->d_q, d_omega);
}

int ibpm_model_C(ibpm_model* m, const double* d_omega, double* d_f) {
    if (!m) { set_error("null model"); return -1; }
    if (!m->reg) return 0;
    const ibpm_flow_params& p = m->p;
    CKR(ibpm_flux_from_vorticity(m->ctx, p.nx, p.ny, p.ngrid, m->dx, 0.0, 0.0, d_omega, m->d_q));
    return ibpm_reg_to_boundary(m->reg, m->d_q, d_f);
}

int ibpm_model_compute_flux(ibpm_model* m, const double* d_omega, double* d_q) {
    if (!m) { set_error("null model"); return -1; }
    const ibpm_flow_params& p = m->p;
    return ibpm_flux_from_vorticity(m->ctx, p.nx, p.ny, p.ngrid, m->dx, m->u0, m->v0, d_omega, d_q);
}

int ibpm_model_constraints(ibpm_model* m, double* h_b) {
    if (!m || (!h_b && m->nb)) { set_error("bad model arguments"); return -1; }
    if (!m->reg) return 0;
    const ibpm_flow_params& p = m->p;
    // q0 = Curl(0) + base flux, then E q0 (getBaseFlowBoundaryVelocities)
    CK(hipMemsetAsync(m->d_w, 0, sizeof(double) * m->ns, m->ctx->c.stream));
    CKR(ibpm_curl_scalar(m->ctx, p.nx, p.ny, p.ngrid, m->dx, m->u0, m->v0, m->d_w, m->d_q));
    CKR(ibpm_reg_to_boundary(m->reg, m->d_q, m->d_y));
    std::vector<double> b0(2 * (size_t)m->nb);
    CK(hipMemcpy(b0.data(), m->d_y, sizeof(double) * b0.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < b0.size(); ++i) h_b[i] = m->uv[i] - b0[i];   // NavierStokesModel.cc:86-91
    return 0;
}

int ibpm_model_M(ibpm_model* m, double beta, const double* d_f, double* d_y) {
    if (!m) { set_error("null model"); return -1; }
    if (!m->reg) return 0;
    const ibpm_flow_params& p = m->p;
    CKR(ibpm_model_B(m, d_f, m->d_w));                                   // B f
    CKR(solve_common(m->ctx, 1, p.nx, p.ny, p.ngrid, m->dx, -beta / 2. / p.reynolds, m->d_w, nullptr, m->d_u,
                     beta));                                              // A^-1 (beta B f)
    return ibpm_model_C(m, m->d_u, d_y);                                  // C A^-1 beta B f
}

int ibpm_cholesky_build(ibpm_model* m, double beta, double* h_M) {
    if (!m) { set_error("null model"); return -1; }
    const int n = 2 * m->nb;
    std::vector<double> e(n, 0.0), col(n);
    for (int j = 0; j < n; ++j) {   // CholeskySolver::computeMatrixM, column j = M e_j
        e.assign(n, 0.0);
        e[j] = 1.0;
        CK(hipMemcpy(m->d_f, e.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        CKR(ibpm_model_M(m, beta, m->d_f, m->d_y));
        CK(hipMemcpy(col.data(), m->d_y, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) h_M[(size_t)i * n + j] = col[i];
    }
    return 0;
}

int ibpm_cholesky_factor(int n, const double* h_M, double* h_lower, double* h_diag) {
    if (n < 0 || (n && (!h_M || !h_lower || !h_diag))) { set_error("bad Cholesky arguments"); return -1; }
    std::vector<double> Mt((size_t)n * n), lower, d;
    for (int r = 0; r < n; ++r)
        for (int c2 = 0; c2 < n; ++c2) Mt[(size_t)c2 * n + r] = h_M[(size_t)r * n + c2];
    const int bad = cholesky_factor(n, Mt, lower, d);
    if (bad >= 0) { set_error("M is not positive definite (pivot %d)", bad); return -2; }
    std::copy(lower.begin(), lower.end(), h_lower);
    std::copy(d.begin(), d.end(), h_diag);
    return 0;
}

int ibpm_cholesky_solve(int n, const double* h_lower, const double* h_diag, const double* h_b, double* h_x) {
    if (n < 0 || (n && (!h_lower || !h_diag || !h_b || !h_x))) { set_error("bad Cholesky arguments"); return -1; }
    // CholeskySolver::Minv (CholeskySolver.cc:200-224); x may alias b
    for (int i = 0; i < n; ++i) {
        double sum = h_b[i];
        for (int k = i - 1; k >= 0; --k) sum -= h_lower[(size_t)i * n + k] * h_x[k];
        h_x[i] = sum / h_diag[i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double sum = h_x[i];
        for (int k = i + 1; k < n; ++k) sum -= h_lower[(size_t)k * n + i] * h_x[k];
        h_x[i] = sum / h_diag[i];
    }
    return 0;
}

int ibpm_cholesky_save(const char* path, int n, double alpha_beta, const double* h_lower, const double* h_diag) {
    if (!path || n < 0) { set_error("bad Cholesky file arguments"); return -1; }
    std::vector<double> lower(h_lower, h_lower + (size_t)n * n), d(h_diag, h_diag + n);
    if (!cholesky_write(path, n, alpha_beta, lower, d)) { set_error("could not write %s", path); return -1; }
    return 0;
}

int ibpm_cholesky_load(const char* path, int n, double alpha_beta, double* h_lower, double* h_diag) {
    if (!path || n < 0) { set_error("bad Cholesky file arguments"); return -1; }
    std::vector<double> lower, d;
    if (!cholesky_read(path, n, alpha_beta, lower, d)) {
        set_error("%s: missing, or not a factor of this size and alpha*beta", path);
        return -2;
    }
    std::copy(lower.begin(), lower.end(), h_lower);
    std::copy(d.begin(), d.end(), h_diag);
    return 0;
}

int ibpm_cg_solve(ibpm_model* m, double beta, double tol, int max_it, const double* h_b, const double* h_x0,
                  double* h_x, int* it_count, int* converged) {
    if (!m || (m->nb && (!h_b || !h_x))) { set_error("bad CG arguments"); return -1; }
    const int n = 2 * m->nb;
    std::vector<double> x(n), r(n), d(n), q(n);
    for (int i = 0; i < n; ++i) x[i] = h_x0 ? h_x0[i] : 0.0;
    auto M = [&](const std::vector<double>& v, std::vector<double>& out) -> int {
        if (n == 0) return 0;
        CK(hipMemcpy(m->d_f, v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        CKR(ibpm_model_M(m, beta, m->d_f, m->d_y));
        CK(hipMemcpy(out.data(), m->d_y, sizeof(double) * n, hipMemcpyDeviceToHost));
        return 0;
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {   // BoundaryVector.h:240-252
        double s2 = 0;
        for (int i = 0; i < n; ++i) s2 += a[i] * b[i];
        return s2;
    };
    // ConjugateGradientSolver::Minv (ConjugateGradientSolver.cc:35-65)
    CKR(M(x, q));
    for (int i = 0; i < n; ++i) r[i] = h_b[i] - q[i];
    d = r;
    double delta = dot(r, r);
    const double tol2 = tol * tol;
    int it = 0;
    while (delta > tol2 && (max_it <= 0 || it < max_it)) {
        CKR(M(d, q));
        const double alpha = delta / dot(d, q);
        for (int i = 0; i < n; ++i) x[i] += alpha * d[i];
        for (int i = 0; i < n; ++i) r[i] -= alpha * q[i];
        const double delta_old = delta;
        delta = dot(r, r);
        const double b2 = delta / delta_old;
        for (int i = 0; i < n; ++i) d[i] = r[i] + b2 * d[i];
        ++it;
    }
    std::copy(x.begin(), x.end(), h_x);
    if (it_count) *it_count = it;
    if (converged) *converged = delta <= tol2 ? 1 : 0;
    return 0;
}

// ------------------------------------------------------------------ batch
int ibpm_batch_create(ibpm_ctx* ctx, int nflows, const ibpm_flow_params* params, ibpm_batch** out) {
    if (!ctx || !params || !out || nflows < 1) { set_error("bad batch arguments"); return -1; }
    CK(hipSetDevice(ctx->c.device));
    auto B = std::make_unique<ibpm_batch>();
    B->ctx = ctx;
    scheme_coefs(params[0].scheme, B.get());
    B->flows.resize(nflows);
    for (int k = 0; k < nflows; ++k) {
        Flow& F = B->flows[k];
        F.p = params[k];
        if (params[k].scheme != params[0].scheme) { set_error("all flows of a batch share one scheme"); return -1; }
        CKR(check_grid(F.p.nx, F.p.ny, F.p.ngrid));
        F.geom_text = F.p.geom ? F.p.geom : "";
        F.p.geom = nullptr;
        std::string err;
        if (!F.geom.parse(F.geom_text, err)) { set_error("flow %d: %s", k, err.c_str()); return -1; }
        F.nx = F.p.nx; F.ny = F.p.ny; F.L = F.p.ngrid;
        F.dx = F.p.length / F.p.nx;   // Grid.cc:77
        F.ns = (long)(F.nx - 1) * (F.ny - 1);
        F.nf = 2L * F.nx * F.ny + F.nx + F.ny;
        F.nb = F.geom.npoints;
        {
            const double pi = 4. * std::atan(1.);
            F.force_alpha = F.p.alpha_deg * pi / 180.;   // ibpm.cc:162-163
        }
        if (F.p.ubf) {   // ibpm.cc:166-178
            if (F.geom.bodies.size() != 1) {
                set_error("flow %d: unsteady base flow is only supported for a single moving body", k);
                return -1;
            }
            if (!F.geom.stationary) {   // Geometry::transferMotion / transferCenter
                Body& b0 = F.geom.bodies[0];
                F.ubf = true;
                F.bf.motion = b0;
                F.bf.motion.rx.clear();
                F.bf.motion.ry.clear();
                F.bf.xc = b0.xc;
                F.bf.yc = b0.yc;
                F.bf.alpha_bf = F.bf.alpha = F.force_alpha;   // BaseFlow(grid, 1, alpha)
                F.bf.mag_bf = F.bf.mag = 1.0;
                b0.motion = M_NONE;
                F.geom.stationary = true;
                B->any_ubf = true;
            }
        }
        if (F.p.model < 0 || F.p.model > 4) { set_error("flow %d: model %d is not 0..4", k, F.p.model); return -1; }
        if (F.p.model == 3 && (F.p.period < 1 || F.p.period > 4096)) {
            set_error("flow %d: linear periodic model needs a period of 1..4096 steps", k);
            return -1;
        }
        if (F.p.model == 4 && !(F.p.sfd_delta > 0)) { set_error("flow %d: SFD needs Delta > 0", k); return -1; }
        if (F.p.model >= 1 && F.p.model <= 3) {   // NavierStokesModel(grid, geom, Re): no potential flow
            if (F.p.ubf) { set_error("flow %d: the linear models take no unsteady base flow", k); return -1; }
            F.base_mag = 0.0;
        }
        if (F.p.model == 3) B->any_periodic = true;
        if (F.p.model == 2) B->any_adjoint = true;
        F.cg = !F.geom.stationary || F.p.force_cg;
        if (F.cg) B->any_cg = true;
        if (!F.geom.stationary) B->all_stationary = false;
        if (F.p.cg_tol <= 0) F.p.cg_tol = 1e-7;   // IBSolver.cc:31 default
        CKR(alloc_flow(B.get(), F));
        CKR(B->mem.alloc(&F.bf_dev, 4));
        {
            const double zero[4] = {0, 0, 0, 0};
            CKR(upload(F.bf_dev, zero, sizeof zero));
        }
        B->max_ns_all = std::max(B->max_ns_all, F.ns * F.L);
        B->max_nf = std::max(B->max_nf, F.nf);
        B->max_cells_all = std::max(B->max_cells_all, (long)(F.nx + 1) * (F.ny + 1) * F.L);
        B->max_comp = std::max(B->max_comp, 2 * F.nb);
        B->max_n = std::max(B->max_n, 2 * F.nb);
        // algorithmic bytes: vel reads omega + q, writes U, V; cross reads U, V
        // and writes the cross flux; rhs reads cross, omega, N_prev and writes
        // a, N; curl_q reads psi and writes q.  vel (and rhs, for the
        // nonlinear model) pass over the nodes strictly inside the coarse
        // levels' central regions, so their bytes count the other nodes, each
        // with its share of the level's fluxes
        const double ncent = (double)(F.nx / 2 - 1) * (F.ny / 2 - 1) * (F.L - 1);
        const double nact = (double)F.ns * F.L - ncent;
        const double fpn = (double)F.nf / F.ns;
        B->bytes_vel += 8.0 * (3 + fpn) * nact;
        B->bytes_cross += 8.0 * (2 * F.ns + F.nf) * F.L;
        B->bytes_rhs += 8.0 * (4 + fpn) * (F.p.model == 0 ? nact : (double)F.ns * F.L);
        B->bytes_curlq += 8.0 * (F.ns + F.nf) * F.L;
    }
    if (B->any_cg) {
        for (auto& F : B->flows) F.cg = true;   // one force solver per batch
    }
    B->max_comp = std::max(B->max_comp, 1);
    B->max_n = std::max(B->max_n, 1);
    CKR(build_plans(B.get()));
    if (B->any_adjoint) {   // the adjoint N's Flux x Flux cross product and its in-place coarsify
        std::vector<CrossFFOps> ops;
        int lmax = 0;
        for (auto& F : B->flows)
            if (F.p.model == 2) {
                ops.push_back(CrossFFOps{F.nx, F.ny, F.L, F.dx, F.bq[0], F.q, F.sff});
                lmax = std::max(lmax, F.L);
            }
        B->cff_n = (int)ops.size();
        CKR(B->mem.alloc(&B->d_cff, ops.size()));
        CKR(upload(B->d_cff, ops.data(), sizeof(CrossFFOps) * ops.size()));
        B->d_cffc.assign(lmax, nullptr);
        B->cffc_n.assign(lmax, 0);
        int rows = 0;
        for (int lev = 1; lev < lmax; ++lev) {
            std::vector<CoarsifyJob> jobs;
            for (auto& F : B->flows)
                if (F.p.model == 2 && lev < F.L) {
                    jobs.push_back(CoarsifyJob{F.nx, F.ny, lev, F.sff, F.sff, nullptr});
                    rows = std::max(rows, F.nx / 2 - 1);
                }
            B->cffc_n[lev] = (int)jobs.size();
            if (jobs.empty()) continue;
            CKR(B->mem.alloc(&B->d_cffc[lev], jobs.size()));
            CKR(upload(B->d_cffc[lev], jobs.data(), sizeof(CoarsifyJob) * jobs.size()));
        }
        B->cffc_rows = rows;
    }
    if (B->any_cg) {
        ibpm_batch* b = B.get();
        b->ring = 2 * b->nsub + 2;
        CKR(b->mem.alloc(&b->nn_params, kNnParams + 1));
        CKR(b->mem.alloc(&b->nn_m, kNnParams + 1));
        CKR(b->mem.alloc(&b->nn_v, kNnParams + 1));
        CKR(b->mem.alloc(&b->nn_grad, kNnParams + 1));
        CKR(b->mem.alloc(&b->d_nn, nflows));
        // W1 ~ U(-a, a) with a = sqrt(6 / (in + hidden)) from a fixed seed;
        // W2 = 0 so the untrained network reproduces the reference warm start
        std::vector<double> P(kNnParams, 0.0);
        unsigned long long st = 0x9E3779B97F4A7C15ull;
        const double a = std::sqrt(6.0 / (kNnIn + kNnHid));
        for (int q = 0; q < kNnHid * kNnIn; ++q) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            P[q] = a * (2.0 * ((st >> 11) * (1.0 / 9007199254740992.0)) - 1.0);
        }
        CKR(upload(b->nn_params, P.data(), sizeof(double) * kNnParams));
    }
    *out = B.release();
    return 0;
}

// Captured step graphs still alive when the process exits (a batch its
// owner never destroyed) are destroyed by an exit handler that runs before
// the HIP runtime's own teardown (registered after the runtime started, so
// it runs first), instead of being left to that teardown.
namespace {
std::mutex g_graph_mu;
std::set<ibpm_batch*> g_graph_batches;
void destroy_live_graphs() {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (ibpm_batch* b : g_graph_batches) {
        (void)hipStreamSynchronize(b->ctx->c.stream);
        if (b->graph) (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    g_graph_batches.clear();
}
void track_graph(ibpm_batch* b, bool live) {
    static bool registered = false;
    std::lock_guard<std::mutex> lk(g_graph_mu);
    if (live) {
        if (!registered) {
            std::atexit(destroy_live_graphs);
            registered = true;
        }
        g_graph_batches.insert(b);
    } else {
        g_graph_batches.erase(b);
    }
}
}  // namespace

int ibpm_batch_destroy(ibpm_batch* b) {
    if (!b) return 0;
    track_graph(b, false);
    (void)hipStreamSynchronize(b->ctx->c.stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (auto e : b->uev)
        if (e) (void)hipEventDestroy(e);
    if (b->poll_ev) (void)hipEventDestroy(b->poll_ev);
    if (b->pc_stream) {
        (void)hipStreamSynchronize(b->pc_stream);
        (void)hipStreamDestroy(b->pc_stream);
    }
    if (b->pc_ev_in) (void)hipEventDestroy(b->pc_ev_in);
    if (b->pc_ev_out) (void)hipEventDestroy(b->pc_ev_out);
    if (b->h_states) (void)hipHostFree(b->h_states);
    if (b->h_reset) (void)hipHostFree(b->h_reset);
    delete b;
    return 0;
}

int ibpm_batch_init(ibpm_batch* B) {
    hipStream_t s = B->ctx->c.stream;
    for (size_t k = 0; k < B->flows.size(); ++k)
        for (size_t ph = 0; ph < B->flows[k].base_set.size(); ++ph)
            if (!B->flows[k].base_set[ph]) {
                set_error("flow %zu: base state %zu of the linear model not set (ibpm_batch_set_base_state)", k, ph);
                return -1;
            }
    for (auto& F : B->flows) {
        F.geom.move(B->time, F.px, F.py, F.pu, F.pv);   // ibpm.cc:285
        CKR(update_regularizer(B, F));                   // model->init()
        if (!F.px.empty()) {
            F.bbox0[0] = *std::min_element(F.px.begin(), F.px.end());
            F.bbox0[1] = *std::max_element(F.px.begin(), F.px.end());
            F.bbox0[2] = *std::min_element(F.py.begin(), F.py.end());
            F.bbox0[3] = *std::max_element(F.py.begin(), F.py.end());
        }
    }
    CKR(upload_regops(B));
    if (!B->any_cg) CKR(build_cholesky(B));              // solver->init()
    for (auto& F : B->flows)                             // updateOperators(t0), ibpm.cc:304
        if (F.ubf) CKR(move_base_flow(B, F, B->time));
    enqueue_refresh(B, s);                               // ibpm.cc:304-305
    CK(hipStreamSynchronize(s));
    CK(hipGetLastError());
    B->initialized = true;
    return 0;
}

// LinearizedPeriodicIBSolver::N (IBSolver.cc:224-231): the base state of phase
// timestep % period for every substep of the coming step.
static int select_phase(ibpm_batch* B) {
    for (int i = 0; i < B->nsub; ++i) {
        bool changed = false;
        for (size_t k = 0; k < B->flows.size(); ++k) {
            Flow& F = B->flows[k];
            if (F.p.model != 3) continue;
            const int ph = B->timestep % F.p.period;
            FlowOps& o = B->h_fops[i][k];
            changed = changed || o.q0b != F.bq[ph];
            o.q0b = F.bq[ph];
            o.w0b = F.bw[ph];
        }
        if (changed) CKR(upload(B->d_fops[i], B->h_fops[i].data(), sizeof(FlowOps) * B->h_fops[i].size()));
    }
    return 0;
}

int ibpm_batch_advance(ibpm_batch* B, int nsteps) {
    if (!B->initialized) { set_error("ibpm_batch_init not called"); return -1; }
    hipStream_t s = B->ctx->c.stream;
    for (int step = 0; step < nsteps; ++step) {
        if (B->any_periodic) CKR(select_phase(B));
        for (auto& F : B->flows)
            if (F.hist_cap <= B->timestep) { set_error("force history capacity exceeded"); return -1; }
        if (B->any_cg) {
            for (int i = 0; i < B->nsub; ++i) {
                CKR(substep_cg(B, i));
                if (i == 0) mark_nprev_valid(B);
            }
            timed(B->probe, K_FORCES, 0, s, [&] { launch_forces(B->d_force, (int)B->flows.size(), s); });
        } else if (B->use_graph && B->old_saved && !B->probe && !B->any_ubf && !B->any_periodic) {
            if (!B->graph) {
                hipGraph_t g;
                CK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                enqueue_step_chol(B, s);
                CK(hipStreamEndCapture(s, &g));
                CK(hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0));
                track_graph(B, true);
                CK(hipGraphDestroy(g));
            }
            CK(hipGraphLaunch(B->graph, s));
        } else {
            for (int i = 0; i < B->nsub; ++i) {
                for (auto& F : B->flows)   // updateOperators(t + c_n dt), IBSolver.cc:164-166
                    if (F.ubf) CKR(move_base_flow(B, F, B->time + B->cn[i] * F.p.dt));
                enqueue_substep_chol(B, i, s);
                if (i == 0 && !B->old_saved) {
                    CK(hipStreamSynchronize(s));
                    mark_nprev_valid(B);
                }
            }
            timed(B->probe, K_FORCES, 0, s, [&] { launch_forces(B->d_force, (int)B->flows.size(), s); });
        }
        B->time += B->flows[0].p.dt;
        B->timestep += 1;
        for (auto& F : B->flows) {   // ibpm.cc:347-351: lab-frame lift and drag
            if (F.ubf) {
                F.bf.set_alpha_mag(B->time);
                F.force_alpha = F.bf.alpha;
            }
            F.alpha_hist.push_back(F.force_alpha);
        }
    }
    CK(hipStreamSynchronize(s));
    CK(hipGetLastError());
    return 0;
}

int ibpm_batch_use_graph(ibpm_batch* B, int enable) {
    B->use_graph = enable != 0;
    return 0;
}

static Flow* flow_of(ibpm_batch* B, int k) {
    if (!B || k < 0 || k >= (int)B->flows.size()) { set_error("bad flow index %d", k); return nullptr; }
    return &B->flows[k];
}

int ibpm_batch_set_base_state(ibpm_batch* B, int k, int ph, const double* h_omega, const double* h_q) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    if (ph < 0 || ph >= (int)F->bq.size()) {
        set_error("flow %d: no base state %d (model %d, period %d)", k, ph, F->p.model, F->p.period);
        return -1;
    }
    CKR(upload(F->bw[ph], h_omega, sizeof(double) * F->ns * F->L));
    CKR(upload(F->bq[ph], h_q, sizeof(double) * F->nf * F->L));
    F->base_set[ph] = true;
    return 0;
}

int ibpm_batch_get_filtered(ibpm_batch* B, int k, double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    if (!F->what) { set_error("flow %d is not an SFD flow", k); return -1; }
    CK(hipStreamSynchronize(B->ctx->c.stream));
    CK(hipMemcpy(h, F->what, sizeof(double) * F->ns * F->L, hipMemcpyDeviceToHost));
    return 0;
}

int ibpm_batch_set_filtered(ibpm_batch* B, int k, const double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    if (!F->what) { set_error("flow %d is not an SFD flow", k); return -1; }
    CK(hipStreamSynchronize(B->ctx->c.stream));
    CKR(upload(F->what, h, sizeof(double) * F->ns * F->L));
    for (int i = 0; i < B->nsub; ++i) {   // the filtered state is set (SFDSolver::_xhatSaved)
        B->h_fops[i][k].sfd_valid = 1;
        CKR(upload(B->d_fops[i], B->h_fops[i].data(), sizeof(FlowOps) * B->h_fops[i].size()));
    }
    if (B->graph) {
        CK(hipGraphExecDestroy(B->graph));
        B->graph = nullptr;
    }
    return 0;
}


int ibpm_batch_num_points(ibpm_batch* B, int k) {
    Flow* F = flow_of(B, k);
    return F ? F->nb : -1;
}

int ibpm_batch_get_omega(ibpm_batch* B, int k, double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    CK(hipMemcpy(h, F->omega, sizeof(double) * F->ns * F->L, hipMemcpyDeviceToHost));
    return 0;
}

int ibpm_batch_set_omega(ibpm_batch* B, int k, const double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    CK(hipMemcpy(F->omega, h, sizeof(double) * F->ns * F->L, hipMemcpyHostToDevice));
    enqueue_refresh(B, B->ctx->c.stream);
    CK(hipStreamSynchronize(B->ctx->c.stream));
    return 0;
}

int ibpm_batch_set_state(ibpm_batch* B, int k, const double* h_omega, const double* h_f, double time,
                         int timestep) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    if (B->timestep != 0 && B->timestep != timestep) {
        set_error("set_state: the batch has advanced %d steps already", B->timestep);
        return -1;
    }
    cancel_ahead(B);   // the time is set anew
    CK(hipMemcpy(F->omega, h_omega, sizeof(double) * F->ns * F->L, hipMemcpyHostToDevice));
    if (h_f) CK(hipMemcpy(F->f, h_f, sizeof(double) * 2 * F->nb, hipMemcpyHostToDevice));
    B->time = time;
    B->timestep = timestep;
    if (!F->geom.stationary) {   // bodies at the restart time (ibpm.cc:285, 304)
        F->geom.move(B->time, F->px, F->py, F->pu, F->pv);
        CKR(update_regularizer(B, *F));
        CKR(upload_regops(B));
    }
    if (F->ubf) {   // base flow at the restart time (ibpm.cc:304)
        CKR(move_base_flow(B, *F, B->time));
        F->force_alpha = F->bf.alpha;
    }
    enqueue_refresh(B, B->ctx->c.stream);
    CK(hipStreamSynchronize(B->ctx->c.stream));
    return 0;
}

int ibpm_batch_get_q(ibpm_batch* B, int k, double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    CK(hipMemcpy(h, F->q, sizeof(double) * F->nf * F->L, hipMemcpyDeviceToHost));
    return 0;
}

int ibpm_batch_get_f(ibpm_batch* B, int k, double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    CK(hipMemcpy(h, F->f, sizeof(double) * 2 * F->nb, hipMemcpyDeviceToHost));
    return 0;
}

int ibpm_batch_get_points(ibpm_batch* B, int k, double* h) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    for (int i = 0; i < F->nb; ++i) {
        h[i] = F->px[i];
        h[F->nb + i] = F->py[i];
    }
    return 0;
}

// Cd, Cl from the net force rotated by the base-flow angle al (radians):
// the free-stream angle, or for an unsteady base flow BaseFlow::getAlpha after
// setAlphaMag at the step's end time.
static void forces_from(const Flow& F, double al, double xf, double yf, double* cd, double* cl) {
    const double dx2 = F.dx * F.dx;   // State.cc:60-62
    xf *= dx2;
    yf *= dx2;
    double sal, cal;
    sin_cos(al, sal, cal);
    const double drag = xf * cal + yf * sal;   // ibpm.cc:352-353
    const double lift = xf * -1. * sal + yf * cal;
    *cd = 2 * drag;
    *cl = 2 * lift;
}

int ibpm_batch_get_forces(ibpm_batch* B, int k, double* cd, double* cl) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    std::vector<double> f(2 * F->nb);
    CK(hipMemcpy(f.data(), F->f, sizeof(double) * f.size(), hipMemcpyDeviceToHost));
    double xf = 0, yf = 0;
    for (int i = 0; i < F->nb; ++i) { xf += f[i]; yf += f[F->nb + i]; }
    forces_from(*F, F->force_alpha, xf, yf, cd, cl);
    return 0;
}

int ibpm_batch_force_history(ibpm_batch* B, int k, double* h, int max_steps) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    int n = 0;
    CK(hipMemcpy(&n, F->hist_count, sizeof(int), hipMemcpyDeviceToHost));
    n = std::min(n, max_steps);
    std::vector<double> raw(2 * (size_t)std::max(n, 1));
    CK(hipMemcpy(raw.data(), F->hist, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
    for (int s = 0; s < n; ++s) {
        const double al = s < (int)F->alpha_hist.size() ? F->alpha_hist[s] : F->force_alpha;
        forces_from(*F, al, raw[2 * s], raw[2 * s + 1], &h[2 * s], &h[2 * s + 1]);
    }
    return n;
}

int ibpm_batch_base_flow(ibpm_batch* B, int k, double* alpha, double* mag) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    if (alpha) *alpha = F->force_alpha;
    if (mag) *mag = F->ubf ? F->bf.mag : 1.0;
    return F->ubf ? 1 : 0;
}

int ibpm_batch_cg_iters(ibpm_batch* B, int k, int* h, int max_sub) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    const int n = std::min(max_sub, B->nsub);
    for (int i = 0; i < n; ++i) h[i] = F->cg ? F->cg_iters[i] : -1;
    return n;
}

double ibpm_batch_time(ibpm_batch* B) { return B ? B->time : 0.0; }

int ibpm_batch_set_warm_start(ibpm_batch* B, int mode) {
    if (mode < 0 || mode > 2) { set_error("warm start mode must be 0, 1 or 2"); return -1; }
    if (mode == 2 && !B->nn_ready) { set_error("learned warm start requested before training"); return -1; }
    B->warm_mode = mode;
    return 0;
}

int ibpm_batch_cholesky_cache(ibpm_batch* B, int k, const char* basename) {
    Flow* F = flow_of(B, k);
    if (!F) return -1;
    if (B->initialized) { set_error("the factor cache must be named before ibpm_batch_init"); return -1; }
    F->chol_cache = basename ? basename : "";
    return 0;
}

int ibpm_batch_cholesky_loaded(ibpm_batch* B, int k) {
    Flow* F = flow_of(B, k);
    return F ? F->chol_loaded : -1;
}

int ibpm_batch_set_precond(ibpm_batch* B, int mode, int pairs) {
    if (!B) { set_error("null batch"); return -1; }
    if (mode < 0 || mode > 1) { set_error("preconditioner mode must be 0 or 1"); return -1; }
    if (pairs < 0 || pairs > kLmpMax) { set_error("pairs must be in 0..%d", kLmpMax); return -1; }
    if (mode == 1 && !B->any_cg) { set_error("preconditioning needs a CG (moving-body) batch"); return -1; }
    B->pc_mode = mode;
    B->pc_pairs = mode ? pairs : 0;
    if (mode) CKR(setup_pcg(B));
    return 0;
}

int ibpm_batch_set_precond_build(ibpm_batch* B, int levels, double rebuild_cost) {
    if (!B) { set_error("null batch"); return -1; }
    if (levels < 0 || levels > 3) {
        set_error("preconditioner build must be 0 (all levels), 1 (finest level), 2 (kernel table) or 3 (kernel "
                  "table, refreshed every substep)");
        return -1;
    }
    if (!(rebuild_cost > 0.0)) { set_error("rebuild cost must be positive"); return -1; }
    if (levels == 1 && !B->helm_allowed) {
        set_error("the finest-level build needs the local Helmholtz path of the force solve");
        return -1;
    }
    B->pc_build_levels = levels;
    B->pc_build_cost = rebuild_cost;
    B->pc_valid = false;   // the next solve builds P the chosen way
    return 0;
}

int ibpm_batch_set_precision(ibpm_batch* B, int mode) {
    if (!B || mode < 0 || mode > 4) { set_error("precision mode %d: 0..4", mode); return -1; }
    B->precision = mode;
    const bool step32 = mode >= 1, fs32 = mode == 2;
    for (int i = 0; i < B->nsub; ++i) {
        set_plan_fp32(B->S1[i], step32);
        set_plan_fp32(B->S2[i], step32);
        set_plan_fp32(B->S3[i], step32);
        set_plan_fp32(B->M1[i], fs32);
    }
    set_plan_fp32(B->S4, step32);
    if (mode >= 3) {   // the step's solves: FP32 radix or split-BF16 GEMM-DST per group
        Ctx* c = &B->ctx->c;
        for (int i = 0; i < B->nsub; ++i) {
            CKR(tune_plan_gemm(c, B->mem, B->S1[i], mode));
            CKR(tune_plan_gemm(c, B->mem, B->S2[i], mode));
            CKR(tune_plan_gemm(c, B->mem, B->S3[i], mode));
        }
        CKR(tune_plan_gemm(c, B->mem, B->S4, mode));
    }
    set_plan_fp32(B->M2, fs32);
    if (!B->M2L.steps.empty()) set_plan_fp32(B->M2L, fs32);
    B->fs32 = fs32;
    if (B->graph) {   // the captured step holds the old kernels
        CK(hipGraphExecDestroy(B->graph));
        B->graph = nullptr;
    }
    return 0;
}

int ibpm_batch_precision_report(ibpm_batch* B, int* n, int* use_gemm, double* t_radix_us, double* t_gemm_us,
                                int max) {
    if (!B) { set_error("null batch"); return -1; }
    std::map<int, const PassGroup*> seen;   // by N: the first group of the step's solves that was timed
    auto visit = [&](const SolvePlan& P) {
        for (auto& step : P.steps)
            for (auto& G : step)
                if (G.gemm_n && (!seen.count(G.gemm_n) || (seen[G.gemm_n]->t_gemm_us == 0 && G.t_gemm_us > 0)))
                    seen[G.gemm_n] = &G;
    };
    for (int i = 0; i < B->nsub; ++i) {
        visit(B->S1[i]);
        visit(B->S2[i]);
        visit(B->S3[i]);
    }
    visit(B->S4);
    int k = 0;
    for (auto& kv : seen) {
        if (k >= max) break;
        if (n) n[k] = kv.first;
        if (use_gemm) use_gemm[k] = kv.second->gemm ? 1 : 0;
        if (t_radix_us) t_radix_us[k] = kv.second->t_radix_us;
        if (t_gemm_us) t_gemm_us[k] = kv.second->t_gemm_us;
        ++k;
    }
    return k;
}

int ibpm_batch_set_local_helmholtz(ibpm_batch* B, int on) {
    if (!B) { set_error("null batch"); return -1; }
    B->helm_allowed = on != 0;
    if (!on)
        for (bool& h : B->helm_on) h = false;
    return 0;
}

int ibpm_batch_precond_matrices(ibpm_batch* B, int flow, int substep, double* h_mtable, double* h_p) {
    if (!B) { set_error("null batch"); return -1; }
    if (flow < 0 || flow >= (int)B->flows.size() || substep < 0 || substep >= B->nsub) {
        set_error("flow %d or substep %d out of range", flow, substep);
        return -1;
    }
    Flow& F = B->flows[flow];
    if (F.geom.stationary || F.nb == 0) { set_error("flow %d has no moving points", flow); return -1; }
    const int n = 2 * F.nb;
    hipStream_t s = B->ctx->c.stream;
    CK(hipStreamSynchronize(s));
    if (B->pc_stream) CK(hipStreamSynchronize(B->pc_stream));
    if (h_mtable) {
        CKR(kernel_tables(B));
        if (!B->d_pcasm_b) CKR(B->mem.alloc(&B->d_pcasm_b, B->flows.size()));
        const PcAsmOps op{n, F.nx, F.ny, F.nsrc, F.ksrc, F.e_sep, F.ktab[substep], beta_of(B, F, substep), F.mcol};
        CKR(upload(B->d_pcasm_b, &op, sizeof op));
        launch_pc_assemble(B->d_pcasm_b, 1, n, s);
        CK(hipStreamSynchronize(s));
        CK(hipGetLastError());
        CK(hipMemcpy(h_mtable, F.mcol, sizeof(double) * n * n, hipMemcpyDeviceToHost));
    }
    if (h_p) {
        if (!B->pc_mode || !B->pc_valid || !F.pinv) { set_error("no preconditioner built"); return -1; }
        CK(hipMemcpy(h_p, F.pinv, sizeof(double) * n * n, hipMemcpyDeviceToHost));
        const double sc = F.pbeta / beta_of(B, F, substep);
        for (long k = 0; k < (long)n * n; ++k) h_p[k] *= sc;
    }
    return 0;
}

int ibpm_batch_solver_stats(ibpm_batch* B, long* stats, int nstats) {
    if (!B || !stats) { set_error("null argument"); return -1; }
    const long v[13] = {B->st_iters, B->st_solves, B->st_mapps, B->st_build_mapps, B->st_builds,
                        B->st_fallbacks, B->pc_broken ? 1L : 0L, (long)(1e6 * B->st_host_kin_s),
                        (long)(1e6 * B->st_host_poll_s), B->st_local_mapps, (long)(1e6 * B->st_build_s),
                        (long)(1e6 * B->st_build_host_s), B->st_refreshes};
    const int n = std::min(nstats, 13);
    for (int k = 0; k < n; ++k) stats[k] = v[k];
    return n;
}

int ibpm_batch_collect(ibpm_batch* B, int enable, long capacity) {
    if (!B->any_cg) { set_error("sample collection needs a CG (moving-body) batch"); return -1; }
    if (enable && capacity > B->data_cap) {
        CKR(B->mem.alloc(&B->data_x, (size_t)capacity * kNnIn));
        CKR(B->mem.alloc(&B->data_y, (size_t)capacity * 2));
        B->data_cap = capacity;
        B->data_n = 0;
    }
    B->collect = enable != 0;
    return 0;
}

long ibpm_batch_num_samples(ibpm_batch* B) { return B ? B->data_n : -1; }

// Full-batch Adam on the collected samples; loss of the last step in *loss.
static int nn_train(Ctx* c, double* P, double* m, double* v, double* grad, long& t, const double* X,
                    const double* Y, long n, int iters, double lr, double* loss) {
    if (n <= 0) { set_error("no training samples"); return -1; }
    DevMem mem;
    double *S = nullptr, *part = nullptr;
    const int nchunks = (int)std::min<long>(256, std::max<long>(1, n / 64));
    CKR(mem.alloc(&S, (size_t)n * kNnScratch));
    CKR(mem.alloc(&part, (size_t)nchunks * (kNnParams + 1)));
    for (int it = 0; it < iters; ++it) {
        double b1t = 1, b2t = 1;
        if (lr > 0) {
            ++t;
            b1t = std::pow(0.9, (double)t);
            b2t = std::pow(0.999, (double)t);
        }
        launch_nn_grad(P, X, Y, n, S, part, nchunks, m, v, grad, lr, b1t, b2t, P, c->stream);
    }
    CK(hipStreamSynchronize(c->stream));
    CK(hipGetLastError());
    if (loss) CK(hipMemcpy(loss, grad + kNnParams, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int ibpm_batch_train(ibpm_batch* B, int iters, double lr, double* loss) {
    if (!B->any_cg) { set_error("training needs a CG (moving-body) batch"); return -1; }
    CKR(nn_train(&B->ctx->c, B->nn_params, B->nn_m, B->nn_v, B->nn_grad, B->nn_adam_t, B->data_x, B->data_y,
                 B->data_n, iters, lr, loss));
    B->nn_ready = true;
    return 0;
}

int ibpm_batch_nn_params(ibpm_batch* B, double* h_get, const double* h_set) {
    if (!B->any_cg) { set_error("no network on a Cholesky batch"); return -1; }
    if (h_set) {
        CKR(upload(B->nn_params, h_set, sizeof(double) * kNnParams));
        B->nn_ready = true;
    }
    if (h_get) CK(hipMemcpy(h_get, B->nn_params, sizeof(double) * kNnParams, hipMemcpyDeviceToHost));
    return kNnParams;
}

int ibpm_nn_loss_grad(ibpm_ctx* c, const double* h_params, const double* h_x, const double* h_y, long n,
                      double* h_grad, double* loss) {
    DevMem mem;
    double *P, *m, *v, *g, *X, *Y;
    CKR(mem.alloc(&P, kNnParams + 1)); CKR(mem.alloc(&m, kNnParams + 1)); CKR(mem.alloc(&v, kNnParams + 1));
    CKR(mem.alloc(&g, kNnParams + 1)); CKR(mem.alloc(&X, (size_t)n * kNnIn)); CKR(mem.alloc(&Y, (size_t)n * 2));
    CKR(upload(P, h_params, sizeof(double) * kNnParams));
    CKR(upload(X, h_x, sizeof(double) * n * kNnIn));
    CKR(upload(Y, h_y, sizeof(double) * n * 2));
    long t = 0;
    CKR(nn_train(&c->c, P, m, v, g, t, X, Y, n, 1, 0.0, loss));
    CK(hipMemcpy(h_grad, g, sizeof(double) * kNnParams, hipMemcpyDeviceToHost));
    return 0;
}

const char* ibpm_kernel_class_name(int cls) {
    return (cls >= 0 && cls < K_NCLASS) ? kclass_names[cls] : "";
}

int ibpm_batch_kernel_stats(ibpm_batch* B, int nsteps, int* counts, double* ms, double* bytes, int nclass) {
    if (!B->initialized) { set_error("ibpm_batch_init not called"); return -1; }
    Probe pr;
    B->probe = &pr;
    const int rc = ibpm_batch_advance(B, nsteps);
    B->probe = nullptr;
    if (rc != 0) return rc;
    for (int k = 0; k < nclass; ++k) { counts[k] = 0; ms[k] = 0; bytes[k] = 0; }
    for (auto& r : pr.recs) {
        if (r.cls >= nclass) continue;
        float t = 0;
        CK(hipEventElapsedTime(&t, r.a, r.b));
        counts[r.cls] += 1;
        ms[r.cls] += t;
        bytes[r.cls] += r.bytes;
    }
    return K_NCLASS;
}

}  // extern "C"
