// omxp_main.cpp -- command-line driver with the reference's surface
// (main.cpp:85-252, :273-736): positional N B [P Q] and the flags -solv,
// -alt, -log, -progress, -init, plus -prec (fp16 | fp8 | learned | fp32), -policy,
// -ir, -seed, -maxit, -restart, -device, --highammat (the Fasi-Higham
// matrix family, main.cpp:239-240, highammgen.hpp:4-30) and -randsvd kappa
// (a conditioned matrix of 2-norm condition number kappa).  One GPU, 1x1
// grid.  With -prec learned the MLP policy picks each step's Schur
// precision; -prec fp32 runs every trailing update in FP32 (panels read
// straight from A, for systems whose kappa * 2^-11 exceeds 1).  With -prec
// fp8 or learned, when the refinement fails the system is generated again
// and factored in FP16 (skipped when the learned schedule was all FP16), and
// if that fails too in FP32, inside the timed region.  -skip is the reference's bit-mask (IR, GETRF,
// TRSM, GEMM; main.cpp:68-70); -progress k prints the reference's
// projected-rate line for every k-th step (panel_trf_gdirect.hpp:1146-1161)
// from per-step completion events, after the factorisation.  The timed region is LU + IR, generation
// excluded (main.cpp:614-657); the rate is (2/3 n^3 + 3/2 n^2) / time
// (main.cpp:663).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <string>
#include <vector>

#include "omxp.h"

namespace {

struct Opts {
    long n = 61440, b = 2560, p = 1, q = 1;
    int solv = 1, alt = 1, log = 1, progress = 10, init = 1, skip = 0;
    int ir = 0, maxit = 50, restart = 50, device = 0;
    int group = 0;  // panels per trailing update: 0 = 4 for fp8/learned, 2 for fp16
    int timer = 0;  // print the phase timers (Timer::show, timer.hpp:160-182) after the run
    // --highammat: 1 = the reference's panel generator, L = I + alpha N and
    // U = I + beta N with N all ones strictly below / above the diagonal
    // (matgen.hpp:157-208); 2 = the paper's form with -alpha, -beta, whose
    // cond_inf is kappa (higham_mat_impl.cpp:130-149)
    int higham = 0;
    double cond = 1e3, rho = 0.5;  // kappa and alpha / beta (higham_mat_impl.cpp:143-148)
    double randsvd = 0;            // -randsvd kappa: A = U diag(sigma) V^T, kappa_2(A) = kappa
    int nrefl = 8;                 // Householder reflectors per orthogonal factor
    double inner_tol = 1e-10;
    unsigned long long seed = 42;
    std::string prec = "fp16";
    std::string policy;
};

void usage() {
    std::printf(
        "usage: omxp N B [P Q] [-solv 0|1] [-alt 0|1|2] [-log 0|1] [-progress k] [-init 1]\n"
        "            [-prec fp16|fp8|learned|fp32] [-policy file] [-ir classic|gmres] [-seed s] [-maxit k]\n"
        "            [-restart m] [-device d] [-group 2|4] [-timer]\n"
        "            [--highammat | --highammat-paper] [-cond kappa] [-rho r]\n"
        "            [-randsvd kappa] [-nrefl k] [-skip mask]\n");
}

bool parse(int argc, char **argv, Opts &o) {
    int pos = 0;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", name);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a[0] != '-') {
            long v = std::atol(a);
            if (pos == 0) o.n = v;
            else if (pos == 1) o.b = v;
            else if (pos == 2) o.p = v;
            else if (pos == 3) o.q = v;
            ++pos;
        } else if (!std::strcmp(a, "-solv")) o.solv = std::atoi(next(a));
        else if (!std::strcmp(a, "-alt")) o.alt = std::atoi(next(a));
        else if (!std::strcmp(a, "-log")) o.log = std::atoi(next(a));
        else if (!std::strcmp(a, "-progress")) o.progress = std::atoi(next(a));
        else if (!std::strcmp(a, "-init")) o.init = std::atoi(next(a));
        else if (!std::strcmp(a, "-skip")) o.skip = std::atoi(next(a));
        else if (!std::strcmp(a, "-prec")) o.prec = next(a);
        else if (!std::strcmp(a, "-policy")) o.policy = next(a);
        else if (!std::strcmp(a, "-ir")) o.ir = !std::strcmp(next(a), "gmres") ? 1 : 0;
        else if (!std::strcmp(a, "-seed")) o.seed = std::strtoull(next(a), nullptr, 10);
        else if (!std::strcmp(a, "-maxit")) o.maxit = std::atoi(next(a));
        else if (!std::strcmp(a, "-restart")) o.restart = std::atoi(next(a));
        else if (!std::strcmp(a, "-device")) o.device = std::atoi(next(a));
        else if (!std::strcmp(a, "-group")) o.group = std::atoi(next(a));
        else if (!std::strcmp(a, "-timer")) o.timer = 1;
        else if (!std::strcmp(a, "--highammat")) o.higham = 1;
        else if (!std::strcmp(a, "--highammat-paper")) o.higham = 2;
        else if (!std::strcmp(a, "-cond")) o.cond = std::atof(next(a));
        else if (!std::strcmp(a, "-rho")) o.rho = std::atof(next(a));
        else if (!std::strcmp(a, "-randsvd")) o.randsvd = std::atof(next(a));
        else if (!std::strcmp(a, "-nrefl")) o.nrefl = std::atoi(next(a));
        else if (!std::strcmp(a, "-h") || !std::strcmp(a, "--help")) {
            usage();
            std::exit(0);
        } else
            std::printf("Ignoring unknown option: %s\n", a);
    }
    return true;
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

}  // namespace

int main(int argc, char **argv) {
    Opts o;
    parse(argc, argv, o);
    if (o.p != 1 || o.q != 1) {
        std::fprintf(stderr, "omxp runs on one GPU: P x Q must be 1 x 1 (got %ld x %ld)\n", o.p, o.q);
        return 2;
    }
    if (o.skip < 0 || o.skip > 15) {
        std::fprintf(stderr, "-skip is a mask of bits 0-3 (IR, GETRF, TRSM, GEMM)\n");
        return 2;
    }
    if (o.init != 1) std::printf("\t-init %d: only the on-device generator (-init 1) exists; using it\n", o.init);
    // the reference's default driver factors the diagonal block on the GPU
    // whatever -solv says (panel_trf_gdirect.hpp:707-711)
    if (o.solv != 1) std::printf("\t-solv %d: the diagonal block is always factored on the GPU; using -solv 1\n", o.solv);
    if (o.randsvd != 0 && (!(o.randsvd >= 1.0) || o.higham || o.nrefl < 1 || o.nrefl > 8)) {
        std::fprintf(stderr, "-randsvd needs kappa >= 1, 1 <= -nrefl <= 8 and no --highammat\n");
        return 2;
    }
    if (o.alt < 0 || o.alt > 2 || o.b <= 0 || o.b % 256) {
        std::fprintf(stderr, "need -alt in 0..2 and B a multiple of 256\n");
        return 2;
    }
    // N rounded to a multiple of B (main.cpp:341-351 with P = Q = 1)
    if (o.n % o.b) {
        long nf = (long)((float)o.n / (float)o.b + 0.5f);
        if (nf == 0) nf = 1;
        o.n = nf * o.b;
    }
    if (o.prec != "fp16" && o.prec != "fp8" && o.prec != "learned" && o.prec != "fp32") {
        std::fprintf(stderr, "-prec must be fp16, fp8, learned or fp32\n");
        return 2;
    }
    if (o.group != 0 && o.group != 2 && o.group != 4) {
        std::fprintf(stderr, "-group must be 2 or 4\n");
        return 2;
    }
    const bool learned = (o.prec == "learned");
    const bool fp8 = (o.prec == "fp8") || learned;
    const bool fp32 = (o.prec == "fp32");
    // the Higham matrix in the library's signed form: a_ij = -beta + ab i
    // above, 1 + ab i on, -alpha + ab j below the diagonal
    double h_alpha = 0, h_beta = 0;
    if (o.higham) {
        const double beta = omx_higham_beta(o.n, o.cond, o.rho);
        if (!std::isfinite(beta) || !(beta > 0)) {
            std::fprintf(stderr, "no Higham beta for n=%ld kappa=%g rho=%g\n", o.n, o.cond, o.rho);
            return 2;
        }
        const double sign = o.higham == 1 ? -1.0 : 1.0;
        h_alpha = sign * o.rho * beta;
        h_beta = sign * beta;
    }
    std::vector<float> params;
    if (learned) {
        if (o.policy.empty()) {
            // next to the package: <pkg>/bin/omxp -> <pkg>/data/sched_mlp.f32
            std::string exe = argv[0];
            const size_t slash = exe.rfind('/');
            const std::string dir = (slash == std::string::npos) ? std::string(".") : exe.substr(0, slash);
            o.policy = dir + "/../data/sched_mlp.f32";
        }
        std::ifstream f(o.policy, std::ios::binary);
        params.resize(OMX_MLP_NPARAMS);
        if (!f.read(reinterpret_cast<char *>(params.data()), OMX_MLP_NPARAMS * sizeof(float))) {
            std::fprintf(stderr, "cannot read %d parameters from %s\n", OMX_MLP_NPARAMS, o.policy.c_str());
            return 2;
        }
    }
    CK(hipSetDevice(o.device));
    const long n = o.n, b = o.b, lda = n;
    std::printf("===============================================================================\n");
    std::printf("omxp - mixed-precision LU + iterative refinement on one MI355X (gfx950)\n");
    std::printf("===============================================================================\n");
    std::printf("\t%s\n", omx_version());
    std::printf("\tn=%ld b=%ld 1x1 grid, alt=%d, Schur precision %s, IR %s (maxit %d)\n", n, b, o.alt,
                learned ? "learned FP8/FP16 per step" : (fp8 ? "FP8 e4m3 (pow2 tile scale)" : (fp32 ? "FP32" : "FP16")),
                o.ir ? "GMRES" : "classical", o.maxit);
    if (o.skip) {
        // main.cpp:440-446
        if (o.skip & OMX_SKIP_IR) std::printf("\tSkipping IR\n");
        if (o.skip & OMX_SKIP_GETRF) std::printf("\tSkipping GETRF\n");
        if (o.skip & OMX_SKIP_TRSM) std::printf("\tSkipping TRSM\n");
        if (o.skip & OMX_SKIP_GEMM) std::printf("\tSkipping GEMM\n");
    }
    if (o.randsvd != 0)
        std::printf("\trandsvd matrix: kappa_2=%g, %d seeded reflectors per orthogonal factor\n", o.randsvd,
                    o.nrefl);
    if (o.higham)
        std::printf("\tFasi-Higham matrix (%s): kappa=%g rho=%g alpha=%.6e beta=%.6e\n",
                    o.higham == 1 ? "L = I + alpha N, U = I + beta N" : "L = I - alpha N, U = I - beta N", o.cond,
                    o.rho, std::fabs(h_alpha), std::fabs(h_beta));

    float *A = nullptr;
    double *d = nullptr, *rhs = nullptr, *x = nullptr;
    CK(hipMalloc(&A, sizeof(float) * (size_t)lda * n));
    CK(hipMalloc(&d, sizeof(double) * n));
    CK(hipMalloc(&rhs, sizeof(double) * n));
    CK(hipMalloc(&x, sizeof(double) * n));
    omx_solver *s = nullptr;
    const int group = o.group ? o.group : (fp8 ? 4 : 2);
    const int flags =
        (fp8 ? OMX_SOLVER_FP8 : 0) | (group == 4 ? OMX_SOLVER_GROUP4 : 0) | (o.timer ? OMX_SOLVER_TIMERS : 0);
    if (omx_solver_create(n, b, o.alt, flags, &s) != 0) {
        std::fprintf(stderr, "solver creation failed\n");
        return 1;
    }
    if (o.restart > OMX_GMRES_RESTART_DEFAULT && o.ir == 1 && omx_solver_reserve_gmres(s, o.restart) != 0) {
        std::fprintf(stderr, "cannot reserve a GMRES basis for restart %d\n", o.restart);
        return 2;
    }
    if (omx_solver_set_skip(s, o.skip) != 0) return 1;
    // randsvd: the factors of diag(sigma) + P Q^T, the FP64 row sums of |A|
    const int rank = 2 * o.nrefl;
    double *sg = nullptr, *P = nullptr, *Q = nullptr, *rowabs = nullptr;
    void *rwork = nullptr;
    size_t rwork_bytes = 0;
    if (o.randsvd != 0) {
        rwork_bytes = omx_randsvd_workspace_bytes(n);
        CK(hipMalloc(&sg, sizeof(double) * n));
        CK(hipMalloc(&P, sizeof(double) * n * rank));
        CK(hipMalloc(&Q, sizeof(double) * n * rank));
        CK(hipMalloc(&rowabs, sizeof(double) * n));
        CK(hipMalloc(&rwork, rwork_bytes));
        if (omx_randsvd_factors(n, o.seed, o.randsvd, o.nrefl, sg, P, Q, nullptr) != 0) {
            std::fprintf(stderr, "randsvd factors failed\n");
            return 1;
        }
    }
    auto generate = [&]() {
        if (o.randsvd != 0) return omx_randsvd_f32(n, rank, sg, P, Q, A, lda, d, rowabs, rwork, rwork_bytes, nullptr);
        return o.higham ? omx_higham_matgen_f32(n, h_alpha, h_beta, A, lda, d, nullptr)
                        : omx_matgen_f32(n, o.seed, A, lda, d, nullptr);
    };
    omx_system sys{};
    sys.kind = o.randsvd != 0 ? OMX_SYSTEM_LOWRANK : (o.higham ? OMX_SYSTEM_HIGHAM : OMX_SYSTEM_LCG);
    sys.seed = o.seed;
    sys.diag = d;
    sys.alpha = h_alpha;
    sys.beta = h_beta;
    sys.rank = rank;
    sys.sigma = sg;
    sys.P = P;
    sys.Q = Q;
    sys.norm_inf = o.higham ? omx_higham_norm_inf(n, h_alpha, h_beta) : 0.0;
    double t = now_s();
    if (generate() || omx_rhs_f64(n, o.seed, rhs, nullptr)) {
        std::fprintf(stderr, "generation failed\n");
        return 1;
    }
    CK(hipDeviceSynchronize());
    if (o.randsvd != 0) {
        // ||A||_inf = max_i sum_j |A_ij| (FP64, from the fill)
        std::vector<double> ra(n);
        CK(hipMemcpy(ra.data(), rowabs, sizeof(double) * n, hipMemcpyDeviceToHost));
        sys.norm_inf = *std::max_element(ra.begin(), ra.end());
    }
    std::printf("\tInitialization = %.4f sec\n", now_s() - t);

    std::vector<int> sched(n / b, fp8 ? 1 : (fp32 ? -1 : 0));
    if (learned && omx_solver_set_policy(s, params.data()) != 0) {
        std::fprintf(stderr, "cannot set the policy\n");
        return 1;
    }
    double tim[5] = {0};
    const double start = now_s();
    int info = omx_lu_factor(s, A, lda, ((fp8 || fp32) && !learned) ? sched.data() : nullptr, tim);
    const double t_lu = now_s() - start;
    if (learned) {
        std::vector<int> used(n / b, 0);
        omx_solver_last_schedule(s, used.data(), nullptr);
        int n8 = 0;
        for (int v : used) n8 += v > 0;
        std::printf("learned schedule: %d of %ld steps in FP8\n", n8, (long)(n / b - 1));
    }
    if (o.progress > 0) {
        // the reference's projected-rate line (panel_trf_gdirect.hpp:1146-1161)
        std::vector<double> ms(n / b, 0.0);
        if (omx_solver_step_times(s, ms.data()) == 0) {
            const double dn = (double)n;
            for (long k = 0; k < n / b; k += o.progress) {
                const double fact = 1.0 - (double)(k * b) / dn;
                const double left = 100.0 * fact * fact * fact;
                const double used_s = ms[k] * 1e-3;
                // as the reference: step 0 leaves 100 % of the work, so its
                // projection is infinite (and its rate 0)
                const double proj_l = used_s / (100.0 - left) * left;
                const double proj_t = used_s + proj_l;
                const double gf = proj_t > 0 ? 2.0 / 3.0 * dn * dn * dn / proj_t * 1.0e-09 : 0.0;
                std::printf("Step %ld, Rank 0, Elapsed %.2lf, Work left %.2lf%%, Time left: %.2lf, Est Total: "
                            "%.2lf, GFlops: %.2lf, TFlops/GPU: %.2lf\n",
                            k, used_s, left, proj_l, proj_t, gf, gf * 0.001);
            }
        }
    }
    std::printf("LU factorization %.2f sec (trailing GEMM %.2f sec over %d launches, %.1f TFLOP/s)\n", t_lu,
                tim[1] * 1e-3, (int)tim[2], tim[1] > 0 ? tim[3] / (tim[1] * 1e-3) * 1e-12 : 0.0);
    if (info) std::printf("LU reported a zero pivot at %d\n", info);
    const double start_ir = now_s();
    double out[4] = {-1.0, -1.0, 0.0, 0.0};
    std::vector<double> trace(2 * o.maxit, 0.0);
    const bool do_ir = !(o.skip & OMX_SKIP_IR);
    int rc = 0;
    if (do_ir) {
        std::printf("\nIR Start with GPU\n");
        rc = omx_ir_solve_system(s, A, lda, &sys, rhs, x, o.ir, o.maxit, o.restart, o.inner_tol, 16.0, 0, out,
                                 trace.data());
    }
    if (do_ir && fp8 && rc != 0) {
        // the fallback ladder: the factors with FP8 Schur steps did not
        // refine; the system again with every step in FP16 (skipped when the
        // learned schedule chose FP16 for every step), then in the FP32 class
        int n8 = 0;
        if (learned) {
            std::vector<int> used(n / b, 0);
            omx_solver_last_schedule(s, used.data(), nullptr);
            for (int v : used) n8 += v > 0;
        }
        const bool try16 = !learned || n8 > 0;
        for (int rung = try16 ? 0 : -1; rung >= -1 && rc != 0; --rung) {
            const char *name = rung == 0 ? "FP16" : "FP32";
            std::printf("schedule did not converge: regenerating and factoring in %s\n", name);
            if (generate()) return 1;
            CK(hipDeviceSynchronize());
            std::fill(sched.begin(), sched.end(), rung);
            info = omx_lu_factor(s, A, lda, sched.data(), tim);
            rc = omx_ir_solve_system(s, A, lda, &sys, rhs, x, o.ir, o.maxit, o.restart, o.inner_tol, 16.0, 0, out,
                                     trace.data());
            std::printf("fallback: %s factorisation, refinement %s\n", name, rc == 0 ? "converged" : "did not converge");
        }
    }
    const double stop = now_s();
    const int steps = !do_ir ? 0 : (rc == 0) ? (int)out[2] + 1 : o.maxit;
    for (int i = 0; i < steps; ++i)
        std::printf("# iterative refinement: step=%d, residual=%e, hpl-harness=%f\n", i, trace[2 * i],
                    trace[2 * i + 1]);
    std::printf("\t IR Time %.2f sec\n", stop - start_ir);
    const double dur = stop - start;
    const double flops = 2.0 / 3.0 * (double)n * n * n + 3.0 / 2.0 * (double)n * n;
    const double resid = (do_ir && rc == 0) ? out[0] : -1.0, harness = (do_ir && rc == 0) ? out[1] : -1.0;
    std::printf("%.3f sec. %.3f GFlop/s resid = %20.15e hpl-harness = %.9f TFlop/s per GPU = %.3f\n", dur,
                flops / dur * 1e-9, resid, harness, flops / dur * 1e-12);
    std::printf("===================================\n");
    std::printf("Code   : omxp\n");
    std::printf("N      : %ld\nNB     : %ld\nP      : 1\nQ      : 1\n", n, b);
    std::printf("PFACT  : Right\nDEPTH  : 1\nPREC   : %s\n", o.prec.c_str());
    std::printf("TIME(s): %.3f\nGFLOPS : %.1f\nRESID  : %20.15e\nHARNESS: %.9f\n", dur, flops / dur * 1e-9, resid,
                harness);
    if (o.timer) {
        // the reference's Timer::show lines (timer.hpp:160-182), GPU seconds per phase
        double sec[OMX_TIMER_ITEMS], ops[OMX_TIMER_ITEMS];
        if (omx_solver_timers(s, sec, ops) == 0) {
            const double total = sec[OMX_TIMER_ITEMS - 2];
            for (int i = 0; i < OMX_TIMER_ITEMS; ++i)
                std::printf(" %-12s : %e sec : %6.2f %% : %20lld : %e Gop/s\n", omx_timer_name(i), sec[i],
                            total > 0 ? 100.0 * sec[i] / total : 0.0, (long long)ops[i],
                            sec[i] > 0 ? 1e-9 * ops[i] / sec[i] : 0.0);
            const int gu = 10, gp = 11;  // GEMM_UPDATE, GEMM_PROGRESS
            const double gt = sec[gu] + sec[gp];
            std::printf("GEMM_TOTAL: %f Tflops\n", gt > 0 ? 1e-12 * (ops[gu] + ops[gp]) / gt : 0.0);
            if (sec[OMX_TIMER_ITEMS - 1] < 0)
                std::printf("(MISC = TOTAL - the items is negative: the trailing updates on the bulk stream run "
                            "beside the factorisation's items, %.3f s of overlap)\n",
                            -sec[OMX_TIMER_ITEMS - 1]);
        }
    }
    omx_solver_destroy(s);
    for (void *pp : {(void *)sg, (void *)P, (void *)Q, (void *)rowabs, rwork})
        if (pp) (void)hipFree(pp);
    (void)hipFree(A);
    (void)hipFree(d);
    (void)hipFree(rhs);
    (void)hipFree(x);
    return rc == 0 ? 0 : 3;
}
