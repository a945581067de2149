// Host-side engine state shared by the staged and fused pipelines.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../include/isce2_amd.h"
#include "fft_lds.h"

namespace ampcor {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::ampcor::Error(AMPCOR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer owned by the engine (grown on demand, kept across calls).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n);
    ~DevBuf();
    template <class U> U *as() const { return static_cast<U *>(ptr); }
};

// FFT plan: radix schedule + FP32 and FP64 twiddle tables on the device.
struct FftPlan {
    int n = 0;
    FftDesc desc32{}, desc64{};
    DevBuf tw32, tw64;
    const void *tw(float) const { return tw32.ptr; }
    const void *tw(double) const { return tw64.ptr; }
};

struct Sizes {
    ampcor_sizes s{};
};

// Device tables the fused kernels read (FP32 twiddles, half-sample shifts).
struct FusedTables {
    const void *tw1, *twW, *twZ, *tw2, *tw256, *h_ref, *h_zoom;
    const void *tw256_64;          // FP64 e^{-2 pi i k/256} (peak refinement)
};

// Pointers of one fused launch: images and starts, per-window stage outputs,
// final outputs (all indexed by absolute window) and the oversampled-chip
// intermediates (this launch's windows only).
struct FusedLaunch {
    const void *ref; int64_t ref_pitch;
    const void *sec; int64_t sec_pitch;
    const int32_t *ref_d, *ref_a, *sec_d, *sec_a;
    int32_t *loc, *start, *shift, *zidx;
    float *val, *mc, *mf;
    float *offset, *snr, *cov;
    float *tamp; double *tmean; float *samp;
    float *surf = nullptr;         // [n][16][16] trimmed surfaces for the refinement (this launch), or null
    double *refined = nullptr;     // [n][2] refined offsets (absolute), or null
};

// Address of this device's copy of the allocation-free entry's scratch array
// (taken once per device, with that device current).
void *noalloc_scratch_init(int device);

struct Engine {
    int device = 0;
    ampcor_params p{};
    ampcor_sizes sz{};
    bool configured = false;
    std::string last_error;
    std::map<int, std::unique_ptr<FftPlan>> plans;
    std::map<std::string, std::unique_ptr<DevBuf>> bufs;
    DevBuf sinc_table;
    int sinc_dec = 4096, sinc_len = 8;

    // per-kernel event timing (ampcor_profile)
    struct ProfRec { std::string name; hipEvent_t a, b; int windows; };
    bool profiling = false;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> event_pool;
    struct ProfAcc { long launches = 0; double ms = 0.0; long windows = 0; };
    std::map<std::string, ProfAcc> prof_acc;
    hipEvent_t take_event();
    std::string prof_only;   // time only this kernel when not empty
    template <class F> void timed(const char *name, int windows, hipStream_t st, F &&launch) {
        if (!profiling || (!prof_only.empty() && prof_only != name)) { launch(); return; }
        ProfRec r{name, take_event(), take_event(), windows};
        (void)hipEventRecord(r.a, st);
        launch();
        (void)hipEventRecord(r.b, st);
        prof_pending.push_back(r);
    }
    void prof_collect();
    ~Engine();

    // set while a call runs on a stream under HIP graph capture: anything that
    // would allocate device memory or copy synchronously throws instead
    bool no_alloc = false;
    // set once a stream capture has recorded launches on this context: a captured
    // graph holds the addresses of its buffers, so none may move (grow) afterwards
    bool captured = false;
    void forbid_in_nosync(const char *what) const {
        if (no_alloc)
            throw Error(AMPCOR_E_ARG, std::string("a call under stream capture would ") + what +
                                          "; run the same call once before capturing it");
    }
    const FftPlan &plan(int n);
    void set_fft_radices(int n, const int32_t *radices, int count);
    template <class U> U *buf(const std::string &name, size_t count) {
        auto &b = bufs[name];
        if (!b) { forbid_in_nosync("allocate a device buffer"); b.reset(new DevBuf()); }
        if (b->bytes < count * sizeof(U)) {
            forbid_in_nosync("grow a device buffer");
            if (captured)
                throw Error(AMPCOR_E_ARG, "a graph captured from this context holds its buffers, and this call "
                                          "needs a larger one: run it on another context");
        }
        b->reserve(count * sizeof(U));
        return b->as<U>();
    }
    void setup(const ampcor_params &params);
    void ensure_sinc_table();

    struct RunArgs {
        const void *ref; int64_t ref_h, ref_w, ref_pitch;
        const void *sec; int64_t sec_h, sec_w, sec_pitch;
        const int32_t *ref_d, *ref_a, *sec_d, *sec_a;
        int32_t n;
        float *offset, *snr, *cov;
        int32_t *maxloc, *start, *shift, *zoom_idx;
        float *maxval;
        int32_t chunk, mode;
        hipStream_t stream;
        float *margin_coarse = nullptr, *margin_fine = nullptr;   // debug: top-1 minus top-2
        bool nosync = false;                                       // AMPCOR_FLAG_NOSYNC
        // windows [fine_from, n) take their coarse zoom start and shift from these
        // (the staged pipeline skips its raw stage for chunks of such windows only)
        const int32_t *fine_start = nullptr, *fine_shift = nullptr;
        int32_t fine_from = 0;
        double *refined = nullptr;   // [n][2] FP64 parabolic refinement of the offset, or null
    };
    // precision: 0 FP32 only; 1 FP32 + FP64 re-run of windows whose argmax
    // margins fall below the thresholds; 2 FP64 staged pipeline for every window
    // FP32 surfaces (values of order 1) carry rounding errors of a few 1e-7;
    // on a 31k-window speckle field every FP32/FP64 flip had a margin <= 6e-8
    // launch-shape options (ampcor_set_option): threads per workgroup of the window-64 fine kernel
    int fine_threads = 1024;
    // window-64 fine kernel: 1 = whole template in LDS (one workgroup per CU),
    // 2 = template streamed in 32-row chunks (two workgroups per CU)
    int fine_kernel = 2;
    // complex elements per workgroup of the staged FFT kernels (float; double uses half)
    int fft_block_elems = 2048;
    // columns per workgroup of the FP32 staged column FFTs (0: from fft_block_elems)
    int fft_cols_width = 8;
    // staged x2 chip oversampling: 1 = quadrant-padded inverse 2-D FFT,
    // 2 = separable half-sample interpolation per axis (equal in exact
    // arithmetic), in-place transforms for the common chip sides; 3 = the
    // separable form with the general out-of-place transforms only
    int oversampler = 2;
    int precision = 1;
    float tau_coarse = 2e-6f, tau_fine = 2e-6f;
    long long prec_checked = 0, prec_rerun = 0;
    void run(const RunArgs &a);
    void run_staged(const RunArgs &a);
    template <class T> void run_staged_t(const RunArgs &a);
    void rerun_fp64(const RunArgs &a, const float *mc, const float *mf, const int32_t *start32 = nullptr,
                    const int32_t *shift32 = nullptr);
    void run_fused(const RunArgs &a);
    FusedTables fused_tables();
    void launch_fused(const FusedTables &tb, const FusedLaunch &c, int n, int w0, hipStream_t st);
    // allocation-free fused run over the device's static scratch array
    void *scratch_ptr = nullptr;
    void run_fused_noalloc(const RunArgs &a);
    void run_noalloc(const RunArgs &a);
};

// FFT schedule for length n (radices in pass order; direct DFT if a prime > 17 remains)
FftDesc make_schedule(int n);
// whether the fused kernels cover a configuration
bool fused_supported(const ampcor_params &p, const ampcor_sizes &s);

}  // namespace ampcor
