// fpm_api.hip -- host side of libfpm_amd: argument checking, buffers, the
// epoch/minibatch loop and the C-ABI entry points declared in include/fpm_amd.h.
//
// The solve follows cuFPM_v3/cuFPM_pure_v3.cu:44-228: per minibatch (a
// contiguous chunk of LEDs) Shat = DFT(O), crop x pupil, batched IDFT, loss
// stencil, batched DFT, back-projection into the spectrum and pupil gradients,
// IDFT of the spectrum gradient, RMSprop of O and P with the pupil clamp;
// lr halves after every epoch with index > 12.  The MI355X schedule differs
// in how the work is grouped, not in what is computed:
//   * the object update is fused with the last inverse pass of the gradient
//     and the first forward pass of the next step's DFT(O) (one HBM pass
//     instead of three), so the spectrum of O is carried from step to step;
//   * the pupil gradient is summed per pixel over the batch in registers and
//     the pupil update runs in the back-projection kernel;
//   * everything is queued on one stream with no host synchronisation inside
//     the solve (the reference synchronises every step, cuFPM_pure_v3.cu:202).
#include "../../include/fpm_amd.h"
#include "common.h"
#include "fft_core.h"
#include "fft_kernels.h"
#include "fpm_kernels.h"
#include "generic.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace fpm {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* get_error() { return g_last_error.c_str(); }

// W_16384^m, m < 16384: FP64-generated, rounded once to FP32; one copy per
// GPU in a static device array (no allocation), filled on the first call that
// needs it on that device
static __device__ __align__(256) float2 g_twiddles[kTwiddleLen];

const float2* twiddle_table(int device) {
    static std::mutex mu;
    static std::map<int, float2*> tables;
    std::lock_guard<std::mutex> lock(mu);
    auto it = tables.find(device);
    if (it != tables.end()) return it->second;
    std::vector<float2> h(kTwiddleLen);
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < kTwiddleLen; ++m) {
        const double a = -two_pi * (double)m / (double)kTwiddleLen;
        h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    void* d = nullptr;
    FPM_HIP(hipGetSymbolAddress(&d, HIP_SYMBOL(g_twiddles)));
    FPM_HIP(hipMemcpy(d, h.data(), sizeof(float2) * kTwiddleLen, hipMemcpyHostToDevice));
    tables[device] = (float2*)d;
    return (const float2*)d;
}

namespace {

struct InvalidArg {
    int code;
    std::string what;
};

[[noreturn]] void invalid(const std::string& w, int code = FPM_E_INVALID) { throw InvalidArg{code, w}; }

// RAII device buffer
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T> T* as() const { return (T*)p; }
    void alloc(size_t bytes) {
        if (bytes == 0) bytes = 16;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw InvalidArg{FPM_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                              hipGetErrorString(e)};
        }
    }
};

// stream: the caller's, or (NULL) one created for the call (default
// priority) that first waits for the work already queued on the null stream
// (torch's default stream, host-side work of a MATLAB session), so inputs
// that work is still producing are read only once they are complete; the
// created stream is non-blocking so that epochs can be captured as graphs
struct StreamHolder {
    hipStream_t s = nullptr;
    bool own = false;
    explicit StreamHolder(void* user) {
        if (user) {
            s = (hipStream_t)user;
            return;
        }
        FPM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        own = true;
        hipEvent_t e = nullptr;
        FPM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        hipError_t r = hipEventRecord(e, nullptr);
        if (r == hipSuccess) r = hipStreamWaitEvent(s, e, 0);
        (void)hipEventDestroy(e);
        FPM_HIP(r);
    }
    ~StreamHolder() {
        if (own && s) (void)hipStreamDestroy(s);
    }
};

fpm_opts merged_opts(const fpm_opts* user) {
    fpm_opts o;
    fpm_opts_init(&o);
    if (user) {
        size_t n = user->struct_size > 0 ? std::min<size_t>((size_t)user->struct_size, sizeof(fpm_opts)) : sizeof(fpm_opts);
        std::memcpy(&o, user, n);
        o.struct_size = sizeof(fpm_opts);
    }
    if (!(o.pupil_clamp > 0.0f)) o.pupil_clamp = 0.2f;
    return o;
}

void check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        invalid("no HIP device available", FPM_E_NODEVICE);
    }
}

void check_side(int64_t v, const char* what) {
    if (!((is_pow2(v) && v >= 16 && v <= 16384) || generic_side(v)))
        invalid(std::string(what) + " = " + std::to_string(v) +
                    ": every image side must be even and in [16, 16384]",
                FPM_E_UNSUPPORTED);
}

Geom make_geom(int64_t Nr, int64_t Nc, int64_t nr, int64_t nc) {
    check_side(Nr, "N_rows");
    check_side(Nc, "N_cols");
    check_side(nr, "n_rows");
    check_side(nc, "n_cols");
    if (nr > Nr || nc > Nc) invalid("the pupil/images are larger than the object");
    Geom g;
    if (!(is_pow2(Nr) && is_pow2(Nc) && is_pow2(nr) && is_pow2(nc))) {
        // a side that is not a power of two: the generic path for all four
        for (int64_t v : {Nr, Nc, nr, nc})
            if (!generic_side(v))
                invalid("with a side that is not a power of two every side must be even and in [16, 16384]",
                        FPM_E_UNSUPPORTED);
        g.nr = (uint32_t)nr;
        g.nc = (uint32_t)nc;
        g.Nr = (uint32_t)Nr;
        g.Nc = (uint32_t)Nc;
        g.lnr = 0;
        g.generic = true;
        ColPerm id{};
        g.pn = id;
        g.pN = id;
        g.pn.M1 = g.nc;
        g.pn.M2 = 1;
        g.pn.C = g.nc;
        g.pN.M1 = g.Nc;
        g.pN.M2 = 1;
        g.pN.C = g.Nc;
        g.nl = colmajor_layout(g.nr, g.pn);
        return g;
    }
    g.nr = (uint32_t)nr;
    g.nc = (uint32_t)nc;
    g.Nr = (uint32_t)Nr;
    g.Nc = (uint32_t)Nc;
    g.lnr = ilog2_host(g.nr);
    g.pn = make_colperm(g.nc);
    g.pN = make_colperm(g.Nc);
    // per-LED pages of a minibatch: 8-row tiles when a column transform fits
    // one pass and the stencil is fused (both sides <= 2048), else column-major
    g.nl = (kTiledPages && g.nr <= 2048 && g.nc <= 2048) ? tiled_layout(g.nc) : colmajor_layout(g.nr, g.pn);
    return g;
}

// device view of a caller array: the pointer itself, or an uploaded copy
template <class T> const T* upload(const T* host_or_dev, size_t count, bool is_device, DevBuf& holder, hipStream_t st) {
    if (is_device) return host_or_dev;
    holder.alloc(sizeof(T) * count);
    FPM_HIP(hipMemcpyAsync(holder.p, host_or_dev, sizeof(T) * count, hipMemcpyHostToDevice, st));
    return holder.as<T>();
}

// image stack on the device, with the uint16 normalisation constants
YStack prepare_y(const void* Y, int32_t y_dtype, size_t count, bool is_device, DevBuf& holder, hipStream_t st) {
    YStack y{};
    if (y_dtype == FPM_Y_F32_AMPLITUDE) {
        y.data = upload((const float*)Y, count, is_device, holder, st);
        y.u16 = 0;
        return y;
    }
    if (y_dtype != FPM_Y_U16_RAW) invalid("y_dtype must be FPM_Y_F32_AMPLITUDE or FPM_Y_U16_RAW");
    const uint16_t* d = upload((const uint16_t*)Y, count, is_device, holder, st);
    DevBuf mm;
    mm.alloc(2 * sizeof(uint32_t));
    const uint32_t init[2] = {0xffffu, 0u};
    FPM_HIP(hipMemcpyAsync(mm.p, init, sizeof(init), hipMemcpyHostToDevice, st));
    launch_u16_minmax(d, count, mm.as<uint32_t>(), st);
    uint32_t h[2];
    FPM_HIP(hipMemcpyAsync(h, mm.p, sizeof(h), hipMemcpyDeviceToHost, st));
    FPM_HIP(hipStreamSynchronize(st));
    if (h[1] <= h[0]) invalid("uint16 image stack is constant (max == min): amplitudes undefined");
    y.data = d;
    y.u16 = 1;
    y.mn = h[0];
    y.range = (float)(h[1] - h[0]);
    y.rcp_range = 1.0f / y.range;
    return y;
}

// the generic path (sides that are not powers of two) runs FP32 mixed-radix
// transforms and the owner-computes stitch only
void check_generic_options(const Geom& g, int eng, int32_t stitch) {
    if (!g.generic) return;
    if (eng == ENG_SPLIT16)
        invalid("the split-FP16 engine needs power-of-two sides (the generic path's transforms are FP32)",
                FPM_E_UNSUPPORTED);
    if (stitch == FPM_STITCH_ATOMIC)
        invalid("the atomic stitch needs power-of-two sides (the generic path stitches owner-computes)",
                FPM_E_UNSUPPORTED);
}

int engine_of(int32_t fft_engine) {
    if (fft_engine == FPM_FFT_F32) return ENG_F32;
    if (fft_engine == FPM_FFT_SPLIT_F16) return ENG_SPLIT16;
    invalid("fft_engine must be FPM_FFT_F32 or FPM_FFT_SPLIT_F16");
}

int run_guarded(const char* fn, const std::function<void()>& body) {
    set_error("");
    try {
        body();
        return FPM_OK;
    } catch (const InvalidArg& e) {
        set_error(std::string(fn) + ": " + e.what);
        return e.code;
    } catch (const HipError& e) {
        set_error(std::string(fn) + ": " + e.what);
        return e.code == hipErrorOutOfMemory ? FPM_E_NOMEM : FPM_E_HIP;
    } catch (const std::exception& e) {
        set_error(std::string(fn) + ": " + e.what());
        return FPM_E_HIP;
    }
}

// Optional per-stage timing with events recorded on the call's stream.
struct StageTimer {
    fpm_stage_times* out = nullptr;
    hipStream_t st = nullptr;
    struct Rec {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    int open_stage = -1;
    hipEvent_t open_ev = nullptr;
    StageTimer(fpm_stage_times* o, hipStream_t s) : out(o), st(s) {
        if (out) std::memset(out, 0, sizeof(*out));
    }
    ~StageTimer() {
        for (auto& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    hipEvent_t ev() {
        hipEvent_t e;
        if (!pool.empty()) {
            e = pool.back();
            pool.pop_back();
        } else {
            FPM_HIP(hipEventCreate(&e));
        }
        return e;
    }
    void begin(int stage) {
        if (!out) return;
        open_stage = stage;
        open_ev = ev();
        FPM_HIP(hipEventRecord(open_ev, st));
    }
    void end(int launches) {
        if (!out) return;
        hipEvent_t e = ev();
        FPM_HIP(hipEventRecord(e, st));
        recs.push_back({open_stage, open_ev, e});
        out->launches[open_stage] = launches;
        // bound the number of live events: fold finished records in
        if (recs.size() >= 512) flush();
    }
    void flush() {
        if (!out) return;
        FPM_HIP(hipStreamSynchronize(st));
        for (auto& r : recs) {
            float ms = 0.f;
            FPM_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            out->ms[r.stage] += ms;
            out->count[r.stage] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
};

// Instantiated epoch graphs, keyed by the caller (a solve: spectrum buffer
// parity and last epoch; the multi-tile solve adds the stream count):
// capture runs the enqueue function on the stream in capture mode (nothing
// executes), then the graph is instantiated once and replayed.
struct GraphCache {
    std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> g;
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() {
        for (auto& kv : g) {
            if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
            if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
        }
    }
    hipGraphExec_t get(int key, hipStream_t st, const std::function<void()>& enqueue) {
        auto it = g.find(key);
        if (it != g.end()) return it->second.second;
        FPM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            enqueue();
        } catch (...) {
            hipGraph_t gr = nullptr;
            (void)hipStreamEndCapture(st, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        hipGraph_t gr = nullptr;
        FPM_HIP(hipStreamEndCapture(st, &gr));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(gr);
            FPM_HIP(e);
        }
        g[key] = {gr, ex};
        return ex;
    }
};

// kernel launches of the column passes of a C-column transform
inline int col_launches(const ColPerm& p) { return p.M2 == 1 ? 1 : 2; }

// Per-minibatch gradient pipeline shared by the solve and the gradient entry.
struct Workspace {
    float2* lat = nullptr;  // b pages each (b <= the batch)
    float2* rec = nullptr;
    DevBuf own_lat, own_rec;
    void alloc(size_t bytes_each) {
        own_lat.alloc(bytes_each);
        own_rec.alloc(bytes_each);
        lat = own_lat.as<float2>();
        rec = own_rec.as<float2>();
    }
};

// Shat (full spectrum, column-permuted) -> per-LED back-projection, the
// pupil part (update or gradient output) and Gs after the column passes of
// its inverse transform.  rowmask (optional): the batch's window rows; the
// column passes then skip the other row tiles of Gs, whose contents are left
// stale and must be read as zero (launch_fused_update does).  pmask
// (optional, nr/16 bytes): the pupil-support row mask of the per-LED spectra
// (launch_pupil_row_mask); with it the per-LED column passes skip pupil rows
// without support (update mode only: the pupil kernel then skips them too).
// pcols (optional): the pupil-support column interval of each pupil row
// (launch_pupil_row_mask); the fused owner stitch then reads P only inside it.
void minibatch_backward(const Geom& g, const float2* Shat, float2* P, const int2* led, uint32_t led0, uint32_t b,
                        const YStack& y, int pratio, Workspace& ws, float2* Gs, int stitch, float* mP, const float* A,
                        RmsArgs rms, float2* GP, double* loss_acc, const float2* tw, hipStream_t st,
                        StageTimer& tm, const uint8_t* rowmask, const uint8_t* pmask, uint8_t* signs = nullptr,
                        const uint32_t* pcols = nullptr) {
    float2* lat = ws.lat;
    float2* rec = ws.rec;
    const bool fused = stencil_fused_supported(g);
    // pupil-support pruning needs the masked read of the fused stencil kernel
    // and the skip of the pupil kernel (update mode, A given)
    const bool prune = pmask && fused && mP && A;
    const uint8_t* lat_mask = prune ? pmask : nullptr;
    const uint8_t* rec_mask = prune ? pmask : nullptr;
    // crop x pupil (kernels.cuh:19-65) is formed inside the first pass
    const CropArgs ca{Shat, P, led, led0, g.nr, g.nc, g.Nr, g.Nc, g.pN, nullptr, 0, pcols};
    if (page_fused_supported(g) && !signs) {
        // 128 x 128 pages: every per-LED stage in one kernel
        tm.begin(FPM_STAGE_STENCIL);
        launch_page_fused(g, ca, rec, y, b, pratio, loss_acc, rec_mask, tw, st, g.eng);
        tm.end(1);
    } else {
        tm.begin(FPM_STAGE_IFFT_N_COLS_LO);
        fft_cols_inverse(lat, g.nr, g.nc, b, g.nl.perm, tw, st, nullptr, &ca, nullptr, lat_mask, &g.nl, g.eng);
        tm.end(col_launches(g.nl.perm));
        if (fused) {
            // contiguous inverse pass + stencil + contiguous forward pass in one kernel
            tm.begin(FPM_STAGE_STENCIL);
            launch_stencil_fused(g, lat, y, led0, b, pratio, rec, loss_acc, lat_mask, tw, st, g.eng, nullptr, 0, signs);
            tm.end(1);
        } else {
            tm.begin(FPM_STAGE_IFFT_N_CONTIG_LO);
            fft_contig(lat, lat, g.nr, g.nc * b, true, tw, st, g.eng);
            tm.end(1);
            tm.begin(FPM_STAGE_STENCIL);
            launch_stencil(g, lat, y, led0, b, pratio, rec, loss_acc, st, signs);
            tm.end(1);
            tm.begin(FPM_STAGE_FFT_N_CONTIG_LO);
            fft_contig(rec, rec, g.nr, g.nc * b, false, tw, st, g.eng);
            tm.end(1);
        }
        tm.begin(FPM_STAGE_FFT_N_COLS_LO);
        // R is read only inside the pupil-support intervals (stitch, pupil update)
        fft_cols_forward(rec, g.nr, g.nc, b, g.nl.perm, tw, st, nullptr, rec_mask, &g.nl, g.eng, 0, nullptr,
                         prune ? pcols : nullptr);
        tm.end(col_launches(g.nl.perm));
    }
    // stitch into Gs, pupil gradient / update, then the column passes of IDFT_N(Gs)
    if (stitch == FPM_STITCH_OWNER && b <= 64) {
        // owner-computes stitch inside the first column pass of IDFT_N; the
        // pupil update must come after it (the stitch reads this step's pupil)
        const StitchArgs sa{rec, P, led, led0, b, g.nr, g.nc, g.Nr, g.Nc, g.nl, 1.0f / (float)(g.Nr * g.Nc), nullptr,
                            pcols};
        tm.begin(FPM_STAGE_IFFT_N_COLS_HI);
        fft_cols_inverse(Gs, g.Nr, g.Nc, 1, g.pN, tw, st, nullptr, nullptr, &sa, rowmask, nullptr, g.eng);
        tm.end(col_launches(g.pN));
        tm.begin(FPM_STAGE_BACKPROJ);
        launch_backproj(g, rec, Shat, led, led0, b, P, Gs, kBackprojPupil, mP, A, rms, GP, st);
        tm.end(1);
        return;
    }
    if (stitch == FPM_STITCH_OWNER) {
        tm.begin(FPM_STAGE_STITCH_OWNER);
        launch_owner_stitch(g, rec, P, led, led0, b, Gs, st, nullptr, 1, pcols);
        tm.end(1);
        tm.begin(FPM_STAGE_BACKPROJ);
        launch_backproj(g, rec, Shat, led, led0, b, P, Gs, kBackprojPupil, mP, A, rms, GP, st);
        tm.end(1);
    } else {
        tm.begin(FPM_STAGE_ZERO);
        launch_zero(Gs, (size_t)g.Nr * g.Nc, st);
        tm.end(1);
        tm.begin(FPM_STAGE_BACKPROJ);
        launch_backproj(g, rec, Shat, led, led0, b, P, Gs, kBackprojAtomic, mP, A, rms, GP, st);
        tm.end(1);
    }
    tm.begin(FPM_STAGE_IFFT_N_COLS_HI);
    fft_cols_inverse(Gs, g.Nr, g.Nc, 1, g.pN, tw, st, nullptr, nullptr, nullptr, rowmask, nullptr, g.eng);
    tm.end(col_launches(g.pN));
}

// One solve's device state and the launches of its epochs: the body of
// fpm_pure_v3, and one per tile of the multi-tile solve.
struct Solve {
    Geom g{};
    YStack y{};
    const int2* led = nullptr;
    const float* A = nullptr;
    float2* O = nullptr;  // working object (= output)
    float2* P = nullptr;  // working pupil (= output)
    uint32_t L = 0, B = 0, steps = 0;
    int pr = 1, stitch = FPM_STITCH_OWNER;
    const float2* tw = nullptr;
    RmsArgs rms{};
    SolveState ss{};
    DevBuf S0, S1, mO, mP, state, masks, ebuf;
    Workspace ws;
    float2* spec[2] = {nullptr, nullptr};
    uint8_t* mk[2] = {nullptr, nullptr};
    uint8_t* pmask = nullptr;  // pupil-support row mask (n side)
    uint8_t* psup = nullptr;   // pupil-support rows
    uint32_t* pcols = nullptr; // pupil-support column interval per pupil row
    uint8_t* csup = nullptr;   // pupil columns inside some row's support interval
    uint8_t* cmask = nullptr;  // window column mask of the next batch (N side)

    void batch_of(uint32_t s, uint32_t& led0, uint32_t& b) const {
        led0 = s * B;
        b = std::min<uint32_t>(B, L - led0);
    }

    // buffers, zeroed moments, device state, row masks and the spectrum of
    // the initial object (rows of the first batch)
    void init(int ne, bool want_loss, float lr, float beta, float eps, float clamp, int32_t halve_from,
              hipStream_t st, StageTimer& tm) {
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        steps = (L + B - 1) / B;
        if (g.generic) {
            init_generic(ne, want_loss, lr, beta, eps, clamp, halve_from, st);
            return;
        }
        S0.alloc(NN * sizeof(float2));
        S1.alloc(NN * sizeof(float2));
        mO.alloc(NN * sizeof(float));
        mP.alloc(nn * sizeof(float));
        ws.alloc(nn * B * sizeof(float2));
        FPM_HIP(hipMemsetAsync(mO.p, 0, NN * sizeof(float), st));
        FPM_HIP(hipMemsetAsync(mP.p, 0, nn * sizeof(float), st));
        // device-resident solve state: lr, epoch counter, FP64 loss slots
        state.alloc(64 + (want_loss ? (size_t)ne * sizeof(double) : 0));
        ss = SolveState{};
        ss.lr = (float*)(state.as<char>() + 8);
        ss.epoch = (uint32_t*)(state.as<char>() + 16);
        ss.loss_cur = want_loss ? (double*)state.p : nullptr;
        ss.loss_hist = want_loss ? (double*)(state.as<char>() + 64) : nullptr;
        ss.halve_from = halve_from > 0 ? (uint32_t)halve_from : 13u;
        launch_solve_state_init(ss, lr, st);
        rms = RmsArgs{lr, beta, eps, clamp, ss.lr};
        // Row masks of the window rows of the current and the next batch:
        // outside them the spectrum gradient is zero and the spectrum is
        // never read, so the N-side column passes skip those row tiles.
        const size_t mrows = g.Nr / kMaskRows, mcols = g.Nc / kMaskRows;
        const size_t cols_at = (2 * mrows + g.nr / kMaskRows + g.nr + mcols + g.nc + 3) & ~(size_t)3;
        masks.alloc(cols_at + g.nr * sizeof(uint32_t));
        pmask = masks.as<uint8_t>() + 2 * mrows;
        psup = pmask + g.nr / kMaskRows;
        cmask = psup + g.nr;
        csup = cmask + mcols;
        pcols = (uint32_t*)(masks.as<uint8_t>() + cols_at);
        mk[0] = masks.as<uint8_t>();
        mk[1] = masks.as<uint8_t>() + mrows;
        spec[0] = S0.as<float2>();
        spec[1] = S1.as<float2>();
        // rows where P != 0 or A != 0: the update keeps P = 0 wherever
        // A = 0, so the masks of the initial pupil hold for every step
        launch_pupil_row_mask(P, A, g.nr, g.nc, pmask, psup, st, pcols);
        launch_pupil_col_support(pcols, g.nr, g.nc, csup, st);
        uint32_t led0, b;
        batch_of(0, led0, b);
        launch_row_mask(led, led0, b, g.Nr, g.nr, psup, mk[0], st);
        // the spectrum is read only inside the batch's windows (rows and
        // columns at the pupil support): its last column pass stores those columns
        launch_row_mask(led, led0, b, g.Nc, g.nc, csup, cmask, st, nullptr, 1, 1);
        tm.begin(FPM_STAGE_SPECTRUM_INIT);
        fft_contig(O, spec[0], g.Nr, g.Nc, false, tw, st, g.eng);
        fft_cols_forward(spec[0], g.Nr, g.Nc, 1, g.pN, tw, st, nullptr, mk[0], nullptr, g.eng, 0, cmask);
        tm.end(1 + col_launches(g.pN));
    }

    // One epoch's launches, given which of the two spectrum buffers (and row
    // masks) holds the current spectrum: every step reads spec[cur] and,
    // unless it is the solve's last step, leaves the next step's spectrum in
    // spec[1 - cur].  The launches depend only on (cur, last): lr, the epoch
    // index and the loss slot live in device memory, so an epoch can be
    // captured and replayed.
    void enqueue_epoch(int cur, bool last_epoch, hipStream_t st, StageTimer& tm) {
        if (g.generic) {
            enqueue_epoch_generic(cur, last_epoch, st, tm);
            return;
        }
        uint32_t led0, b;
        for (uint32_t s = 0; s < steps; ++s) {
            batch_of(s, led0, b);
            minibatch_backward(g, spec[cur], P, led, led0, b, y, pr, ws, spec[1 - cur], stitch, mP.as<float>(), A,
                               rms, nullptr, ss.loss_cur, tw, st, tm, mk[cur], pmask, nullptr, pcols);
            const bool last = last_epoch && (s == steps - 1);
            if (!last) {  // the next minibatch: s + 1, or the first one of the next epoch
                uint32_t nled0, nb;
                batch_of(s + 1 < steps ? s + 1 : 0, nled0, nb);
                launch_row_mask(led, nled0, nb, g.Nr, g.nr, psup, mk[1 - cur], st);
                launch_row_mask(led, nled0, nb, g.Nc, g.nc, csup, cmask, st, nullptr, 1, 1);
            }
            tm.begin(FPM_STAGE_FUSED_UPDATE);
            launch_fused_update(g, spec[1 - cur], O, mO.as<float>(), rms, !last, mk[cur], last ? nullptr : mk[1 - cur],
                                tw, st, g.eng);
            tm.end(1);
            if (!last) {
                tm.begin(FPM_STAGE_FFT_N_COLS_HI);
                fft_cols_forward(spec[1 - cur], g.Nr, g.Nc, 1, g.pN, tw, st, nullptr, mk[1 - cur], nullptr, g.eng, 0,
                                 cmask);
                tm.end(col_launches(g.pN));
                cur = 1 - cur;
            }
        }
        // loss of the epoch to its slot; lr halving (cuFPM_pure_v3.cu:204-206)
        launch_epoch_end(ss, st);
    }

    // spectrum buffer parity at the start of the next epoch
    int parity_after(int cur) const { return (int)((cur + steps) & 1u); }

    // generic sides (generic.hip): S = DFT_N(O) every step into S0, the
    // gradient spectrum and G_O in S1, the residuals in ebuf (unfused pages only)
    void init_generic(int ne, bool want_loss, float lr, float beta, float eps, float clamp, int32_t halve_from,
                      hipStream_t st) {
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        S0.alloc(NN * sizeof(float2));
        S1.alloc(NN * sizeof(float2));
        mO.alloc(NN * sizeof(float));
        mP.alloc(nn * sizeof(float));
        ws.alloc(nn * B * sizeof(float2));
        if (!generic_fused_pages(g, false)) ebuf.alloc(nn * B * sizeof(float));
        FPM_HIP(hipMemsetAsync(mO.p, 0, NN * sizeof(float), st));
        FPM_HIP(hipMemsetAsync(mP.p, 0, nn * sizeof(float), st));
        state.alloc(64 + (want_loss ? (size_t)ne * sizeof(double) : 0));
        ss = SolveState{};
        ss.lr = (float*)(state.as<char>() + 8);
        ss.epoch = (uint32_t*)(state.as<char>() + 16);
        ss.loss_cur = want_loss ? (double*)state.p : nullptr;
        ss.loss_hist = want_loss ? (double*)(state.as<char>() + 64) : nullptr;
        ss.halve_from = halve_from > 0 ? (uint32_t)halve_from : 13u;
        launch_solve_state_init(ss, lr, st);
        rms = RmsArgs{lr, beta, eps, clamp, ss.lr};
        spec[0] = S0.as<float2>();
        spec[1] = S1.as<float2>();
        // pupil-support column intervals (the stitch reads P only inside them)
        masks.alloc(g.nr * sizeof(uint32_t));
        pcols = masks.as<uint32_t>();
        generic_pupil_cols(P, A, g, pcols, st);
        generic_prepare(g);
        generic_tune(g, 1, B, O, S0.as<float2>(), P, led, nullptr, ws.lat, st);
        // with the fused update the spectrum is carried from step to step
        // (S0 / S1 alternate as on the power-of-two path): the first one here
        if (generic_update_fusable(g)) fft2_generic(O, spec[0], g.Nr, g.Nc, 1, false, st);
    }

    // Stage events: each whole 2D transform is recorded under its column-pass
    // stage (FFT_N_COLS_HI: DFT_N(O); IFFT_N_COLS_LO / FFT_N_COLS_LO: the
    // per-LED transforms; IFFT_N_COLS_HI: IDFT_N of the gradient), the
    // residual + stencil under STENCIL and the RMSprop of O under
    // FUSED_UPDATE.
    // With the fused update (N_rows = 3 x 2^k) the step's spectrum comes from
    // the previous step (spec[cur]) and the gradient's strided inverse pass,
    // the fused contiguous inverse + RMSprop + contiguous forward pass and the
    // strided forward pass leave the next one in spec[1 - cur].
    void enqueue_epoch_generic(int cur, bool last_epoch, hipStream_t st, StageTimer& tm) {
        const size_t NN = (size_t)g.Nr * g.Nc;
        const bool fuse_update = generic_update_fusable(g);
        float2* lat = ws.lat;
        float2* rec = ws.rec;
        const bool fused = generic_fused_pages(g, false);
        uint32_t led0, b;
        for (uint32_t s = 0; s < steps; ++s) {
            batch_of(s, led0, b);
            float2* S = fuse_update ? spec[cur] : S0.as<float2>();
            float2* G = fuse_update ? spec[1 - cur] : S1.as<float2>();
            if (!fuse_update) {
                tm.begin(FPM_STAGE_FFT_N_COLS_HI);
                fft2_generic(O, S, g.Nr, g.Nc, 1, false, st);
                tm.end(2);
            }
            if (fused) {
                tm.begin(FPM_STAGE_IFFT_N_COLS_LO);  // crop x pupil inside the row pass
                generic_inverse_pages(S, P, led, led0, b, g, lat, st, nullptr, 1, true);
                tm.end(1);
            } else {
                tm.begin(FPM_STAGE_CROP);
                generic_crop(S, P, led, led0, b, g, lat, st);
                tm.end(1);
                tm.begin(FPM_STAGE_IFFT_N_COLS_LO);
                fft2_generic(lat, lat, g.nr, g.nc, b, true, st);
                tm.end(2);
            }
            tm.begin(FPM_STAGE_STENCIL);
            generic_stencil_pages(lat, y, led0, b, pr, g, ebuf.as<float>(), rec, ss.loss_cur, st, nullptr, 1, nullptr,
                                  fused);
            tm.end(fused ? 1 : 2);
            tm.begin(FPM_STAGE_FFT_N_COLS_LO);
            generic_forward_pages(rec, b, g, st, fused);
            tm.end(fused ? 1 : 2);
            tm.begin(FPM_STAGE_STITCH_OWNER);
            generic_stitch(rec, P, led, led0, b, g, G, st, nullptr, 1, pcols);  // reads this step's pupil
            tm.end(1);
            tm.begin(FPM_STAGE_BACKPROJ);
            generic_pupil(rec, S, led, led0, b, g, P, mP.as<float>(), A, rms, nullptr, st);
            tm.end(1);
            if (fuse_update) {
                const bool last = last_epoch && (s == steps - 1);
                tm.begin(FPM_STAGE_IFFT_N_COLS_HI);
                generic_strided_pass(G, g.Nr, g.Nc, 1, true, st);
                tm.end(1);
                tm.begin(FPM_STAGE_FUSED_UPDATE);
                generic_update_fused(G, O, mO.as<float>(), g, 1, rms, !last, st);
                tm.end(1);
                if (!last) {
                    tm.begin(FPM_STAGE_FFT_N_COLS_HI);
                    generic_strided_pass(G, g.Nr, g.Nc, 1, false, st);
                    tm.end(1);
                    cur = 1 - cur;
                }
                continue;
            }
            tm.begin(FPM_STAGE_IFFT_N_COLS_HI);
            fft2_generic(G, G, g.Nr, g.Nc, 1, true, st);
            tm.end(2);
            tm.begin(FPM_STAGE_FUSED_UPDATE);
            generic_update(G, O, mO.as<float>(), NN, rms, st);
            tm.end(1);
        }
        launch_epoch_end(ss, st);
    }
};

// Tiles of one size and image type run through shared launches: every stage
// of a minibatch step takes all T tiles at once -- T b pages for the per-LED
// stages (tile k's pages k b..), T pages (column passes, stitch) or T N_cols
// lines (fused update) for the object side.  Tile k owns block k of every
// working array (O, P, both spectra, moments, pages, row masks) and reads its
// images, LED windows, pupil mask and loss slot through a device table.  Each
// element is computed with the arithmetic of the one-tile Solve, so every
// tile's result is the same bits as its own solve.
struct GroupSolve {
    Geom g{};
    uint32_t T = 0, L = 0, B = 0, steps = 0, ne = 0;
    int pr = 1;
    const float2* tw = nullptr;
    YStack y{};  // image type; each tile's stack and constants are in the table
    RmsArgs rms{};
    SolveState ss{};
    DevBuf O, P, S0, S1, mO, mP, lat, rec, masks, state, tab, ebuf;
    std::vector<std::unique_ptr<DevBuf>> ybuf;
    std::vector<TileIn> htab;
    float2* spec[2] = {nullptr, nullptr};
    uint8_t* mk[2] = {nullptr, nullptr};
    uint8_t* pmask = nullptr;  // T pupil-support row masks (n side)
    uint8_t* psup = nullptr;   // T pupil-support row sets
    uint32_t* pcols = nullptr; // T pupil-support column intervals (nr words each)
    const TileIn* tabd = nullptr;

    // The owner stitch as its own kernel (coalesced gathers along the rows,
    // every element written) + a plain column pass, or summed inside the
    // column pass (one gather per slot, strided; rows outside the batch's
    // windows skipped).  Same sums in the same LED order: the same bits.
    // Up to 1024^2 objects the separate kernel measured faster (24 tiles of
    // 1024^2 + 24 of 512^2: 25.6 -> 22.3 ms per epoch); at 16384^2 writing the
    // whole gradient costs more than the strided gathers (kept fused there).
    bool separate_stitch() const { return (size_t)g.Nr * g.Nc <= (size_t)1024 * 1024; }

    // tiles of one group can share launches
    static bool groupable(const Geom& g, uint32_t B, int stitch) {
        if (g.generic) return stitch == FPM_STITCH_OWNER;  // the generic kernels take a tile grid row
        return stitch == FPM_STITCH_OWNER && B <= (uint32_t)64 && stencil_fused_supported(g) &&
               fused_update_groupable(g);
    }

    void batch_of(uint32_t s, uint32_t& led0, uint32_t& b) const {
        led0 = s * B;
        b = std::min<uint32_t>(B, L - led0);
    }

    void init(const fpm_tile* tiles, const std::vector<int>& members, int ne_, float lr, float beta, float eps,
              float clamp, int32_t halve_from, hipStream_t st) {
        T = (uint32_t)members.size();
        ne = (uint32_t)ne_;
        if ((uint64_t)T * B > 65535u) invalid("a tile group holds more than 65535 pages per step");
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        steps = (L + B - 1) / B;
        bool want_loss = false;
        for (int t : members) want_loss |= tiles[t].loss_per_epoch != nullptr;
        O.alloc(T * NN * sizeof(float2));
        P.alloc(T * nn * sizeof(float2));
        S0.alloc(T * NN * sizeof(float2));
        S1.alloc(T * NN * sizeof(float2));
        mO.alloc(T * NN * sizeof(float));
        mP.alloc(T * nn * sizeof(float));
        lat.alloc((size_t)T * B * nn * sizeof(float2));
        rec.alloc((size_t)T * B * nn * sizeof(float2));
        const size_t mrows = g.Nr / kMaskRows, prows = g.nr / kMaskRows;
        const size_t cols_at = (T * (2 * mrows + prows + g.nr) + 3) & ~(size_t)3;
        masks.alloc(cols_at + (size_t)T * g.nr * sizeof(uint32_t));
        mk[0] = masks.as<uint8_t>();
        mk[1] = mk[0] + T * mrows;
        pmask = mk[1] + T * mrows;
        psup = pmask + T * prows;
        pcols = (uint32_t*)(masks.as<uint8_t>() + cols_at);
        spec[0] = S0.as<float2>();
        spec[1] = S1.as<float2>();
        state.alloc(64 + (want_loss ? (size_t)T * (1 + ne) * sizeof(double) : 0));
        ss = SolveState{};
        ss.lr = (float*)(state.as<char>() + 8);
        ss.epoch = (uint32_t*)(state.as<char>() + 16);
        ss.loss_cur = want_loss ? (double*)(state.as<char>() + 64) : nullptr;
        ss.loss_hist = want_loss ? ss.loss_cur + T : nullptr;
        ss.halve_from = halve_from > 0 ? (uint32_t)halve_from : 13u;
        ss.ntiles = T;
        ss.hist_stride = ne;
        htab.resize(T);
        // uint16 stacks (on the device): every tile's min / max reduction is
        // launched first and read back once (prepare_y's steps and checks)
        DevBuf mm;
        mm.alloc((size_t)T * 2 * sizeof(uint32_t));
        std::vector<uint32_t> mmh(2 * (size_t)T);
        for (uint32_t k = 0; k < T; ++k) {
            mmh[2 * k] = 0xffffu;
            mmh[2 * k + 1] = 0u;
        }
        FPM_HIP(hipMemcpyAsync(mm.p, mmh.data(), mmh.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        std::vector<YStack> ys(T);
        for (uint32_t k = 0; k < T; ++k) {
            const fpm_tile& tl = tiles[members[k]];
            FPM_HIP(hipMemcpyAsync(O.as<float2>() + k * NN, tl.O0, NN * sizeof(float2), hipMemcpyDeviceToDevice, st));
            FPM_HIP(hipMemcpyAsync(P.as<float2>() + k * nn, tl.P0, nn * sizeof(float2), hipMemcpyDeviceToDevice, st));
            if (tl.y_dtype == FPM_Y_U16_RAW) {
                ys[k].data = tl.Y;
                ys[k].u16 = 1;
                launch_u16_minmax((const uint16_t*)tl.Y, nn * (size_t)L, mm.as<uint32_t>() + 2 * k, st);
            } else {
                ybuf.push_back(std::make_unique<DevBuf>());
                ys[k] = prepare_y(tl.Y, tl.y_dtype, nn * (size_t)L, true, *ybuf.back(), st);
            }
        }
        FPM_HIP(hipMemcpyAsync(mmh.data(), mm.p, mmh.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        FPM_HIP(hipStreamSynchronize(st));
        for (uint32_t k = 0; k < T; ++k) {
            const fpm_tile& tl = tiles[members[k]];
            YStack& yk = ys[k];
            if (yk.u16) {
                if (mmh[2 * k + 1] <= mmh[2 * k])
                    invalid("uint16 image stack is constant (max == min): amplitudes undefined");
                yk.mn = mmh[2 * k];
                yk.range = (float)(mmh[2 * k + 1] - mmh[2 * k]);
                yk.rcp_range = 1.0f / yk.range;
            }
            y.u16 = yk.u16;
            htab[k] = TileIn{(const int2*)tl.led_idx, yk.data, yk.mn, yk.range, yk.rcp_range, 0u, tl.A,
                             (want_loss && tl.loss_per_epoch) ? ss.loss_cur + k : nullptr};
        }
        tab.alloc(T * sizeof(TileIn));
        FPM_HIP(hipMemcpyAsync(tab.p, htab.data(), T * sizeof(TileIn), hipMemcpyHostToDevice, st));
        tabd = tab.as<TileIn>();
        FPM_HIP(hipMemsetAsync(mO.p, 0, T * NN * sizeof(float), st));
        FPM_HIP(hipMemsetAsync(mP.p, 0, T * nn * sizeof(float), st));
        launch_solve_state_init(ss, lr, st);
        rms = RmsArgs{lr, beta, eps, clamp, ss.lr};
        if (g.generic) {  // no row masks; the spectrum is carried only with the fused update
            generic_pupil_cols(P.as<float2>(), nullptr, g, pcols, st, tabd, T);
            if (!generic_fused_pages(g, false)) ebuf.alloc((size_t)T * B * nn * sizeof(float));
            generic_prepare(g);
            generic_tune(g, T, B, O.as<float2>(), S0.as<float2>(), P.as<float2>(), nullptr, tabd, lat.as<float2>(),
                         st);
            if (generic_update_fusable(g)) fft2_generic(O.as<float2>(), spec[0], g.Nr, g.Nc, T, false, st);
            return;
        }
        for (uint32_t k = 0; k < T; ++k)
            launch_pupil_row_mask(P.as<float2>() + k * nn, tiles[members[k]].A, g.nr, g.nc, pmask + k * prows,
                                  psup + (size_t)k * g.nr, st, pcols + (size_t)k * g.nr);
        uint32_t led0, b;
        batch_of(0, led0, b);
        launch_row_mask(nullptr, led0, b, g.Nr, g.nr, psup, mk[0], st, tabd, T);
        fft_contig(O.as<float2>(), spec[0], g.Nr, g.Nc * T, false, tw, st, g.eng);
        fft_cols_forward(spec[0], g.Nr, g.Nc, T, g.pN, tw, st, nullptr, mk[0], nullptr, g.eng, 1);
    }

    // the launches of Solve::enqueue_epoch + minibatch_backward (update
    // mode, fused stencil, owner stitch inside the column pass), for all tiles
    void enqueue_epoch(int cur, bool last_epoch, hipStream_t st) {
        if (g.generic) {
            enqueue_epoch_generic(cur, last_epoch, st);
            return;
        }
        const float invN2 = 1.0f / (float)(g.Nr * g.Nc);
        float2* Pw = P.as<float2>();
        float2* lp = lat.as<float2>();
        float2* rp = rec.as<float2>();
        uint32_t led0, b;
        for (uint32_t s = 0; s < steps; ++s) {
            batch_of(s, led0, b);
            const uint32_t pages = T * b;
            const CropArgs ca{spec[cur], Pw, nullptr, led0, g.nr, g.nc, g.Nr, g.Nc, g.pN, tabd, b, pcols};
            if (page_fused_supported(g)) {
                launch_page_fused(g, ca, rp, y, pages, pr, nullptr, pmask, tw, st, g.eng);
            } else {
                fft_cols_inverse(lp, g.nr, g.nc, pages, g.nl.perm, tw, st, nullptr, &ca, nullptr, pmask, &g.nl, g.eng,
                                 b);
                launch_stencil_fused(g, lp, y, led0, pages, pr, rp, nullptr, pmask, tw, st, g.eng, tabd, b);
                fft_cols_forward(rp, g.nr, g.nc, pages, g.nl.perm, tw, st, nullptr, pmask, &g.nl, g.eng, b, nullptr,
                                 pcols);
            }
            if (separate_stitch()) {
                launch_owner_stitch(g, rp, Pw, nullptr, led0, b, spec[1 - cur], st, tabd, T, pcols);
                fft_cols_inverse(spec[1 - cur], g.Nr, g.Nc, T, g.pN, tw, st, nullptr, nullptr, nullptr, mk[cur], nullptr,
                                 g.eng, 1);
            } else {
                const StitchArgs sa{rp, Pw, nullptr, led0, b, g.nr, g.nc, g.Nr, g.Nc, g.nl, invN2, tabd, pcols};
                fft_cols_inverse(spec[1 - cur], g.Nr, g.Nc, T, g.pN, tw, st, nullptr, nullptr, &sa, mk[cur], nullptr,
                                 g.eng, 1);
            }
            launch_backproj(g, rp, spec[cur], nullptr, led0, b, Pw, spec[1 - cur], kBackprojPupil, mP.as<float>(),
                            nullptr, rms, nullptr, st, tabd, T);
            const bool last = last_epoch && (s == steps - 1);
            if (!last) {
                uint32_t nled0, nb;
                batch_of(s + 1 < steps ? s + 1 : 0, nled0, nb);
                launch_row_mask(nullptr, nled0, nb, g.Nr, g.nr, psup, mk[1 - cur], st, tabd, T);
            }
            launch_fused_update(g, spec[1 - cur], O.as<float2>(), mO.as<float>(), rms, !last, mk[cur],
                                last ? nullptr : mk[1 - cur], tw, st, g.eng,
                                T);
            if (!last) {
                fft_cols_forward(spec[1 - cur], g.Nr, g.Nc, T, g.pN, tw, st, nullptr, mk[1 - cur], nullptr, g.eng, 1);
                cur = 1 - cur;
            }
        }
        launch_epoch_end(ss, st);
    }

    // the generic steps (Solve::enqueue_epoch_generic) for all T tiles: T
    // pages for the object-side transforms, T b for the per-LED ones, one
    // grid row per tile for the crop, stencil, stitch and pupil kernels
    void enqueue_epoch_generic(int cur, bool last_epoch, hipStream_t st) {
        const size_t NN = (size_t)g.Nr * g.Nc;
        const bool fuse_update = generic_update_fusable(g);
        float2* Ow = O.as<float2>();
        float2* Pw = P.as<float2>();
        float2* lp = lat.as<float2>();
        float2* rp = rec.as<float2>();
        const bool fused = generic_fused_pages(g, false);
        uint32_t led0, b;
        for (uint32_t s = 0; s < steps; ++s) {
            batch_of(s, led0, b);
            float2* S = fuse_update ? spec[cur] : S0.as<float2>();
            float2* G = fuse_update ? spec[1 - cur] : S1.as<float2>();
            if (!fuse_update) fft2_generic(Ow, S, g.Nr, g.Nc, T, false, st);
            generic_inverse_pages(S, Pw, nullptr, led0, b, g, lp, st, tabd, T, fused);
            generic_stencil_pages(lp, y, led0, b, pr, g, ebuf.as<float>(), rp, nullptr, st, tabd, T, nullptr, fused);
            generic_forward_pages(rp, T * b, g, st, fused);
            generic_stitch(rp, Pw, nullptr, led0, b, g, G, st, tabd, T, pcols);
            generic_pupil(rp, S, nullptr, led0, b, g, Pw, mP.as<float>(), nullptr, rms, nullptr, st, tabd, T);
            if (fuse_update) {
                const bool last = last_epoch && (s == steps - 1);
                generic_strided_pass(G, g.Nr, g.Nc, T, true, st);
                generic_update_fused(G, Ow, mO.as<float>(), g, T, rms, !last, st);
                if (!last) {
                    generic_strided_pass(G, g.Nr, g.Nc, T, false, st);
                    cur = 1 - cur;
                }
                continue;
            }
            fft2_generic(G, G, g.Nr, g.Nc, T, true, st);
            generic_update(G, Ow, mO.as<float>(), T * NN, rms, st);
        }
        launch_epoch_end(ss, st);
    }

    // results (and loss histories) to the callers' arrays
    void finish(const fpm_tile* tiles, const std::vector<int>& members, int epochs, hipStream_t st) {
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        for (uint32_t k = 0; k < T; ++k) {
            const fpm_tile& tl = tiles[members[k]];
            FPM_HIP(hipMemcpyAsync(tl.O_out, O.as<float2>() + k * NN, NN * sizeof(float2), hipMemcpyDeviceToDevice,
                                   st));
            FPM_HIP(hipMemcpyAsync(tl.P_out, P.as<float2>() + k * nn, nn * sizeof(float2), hipMemcpyDeviceToDevice,
                                   st));
            if (tl.loss_per_epoch && ss.loss_hist)
                FPM_HIP(hipMemcpyAsync(tl.loss_per_epoch, ss.loss_hist + (size_t)k * ne, (size_t)epochs * sizeof(double),
                                       hipMemcpyDeviceToHost, st));
        }
    }

    int parity_after(int cur) const { return (int)((cur + steps) & 1u); }
};

}  // namespace
}  // namespace fpm

using namespace fpm;

extern "C" {

void fpm_opts_init(fpm_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(fpm_opts);
    // owner-computes: measured faster than the atomics at 16384^2 / batch 26
    // on MI355X and deterministic; FPM_STITCH_ATOMIC remains selectable
    o->stitch = FPM_STITCH_OWNER;
    o->pointers_are_device = 0;
    o->verbose = 0;
    o->stream = nullptr;
    o->loss_per_epoch = nullptr;
    o->pupil_clamp = 0.2f;
}

const char* fpm_last_error(void) { return get_error(); }

int32_t fpm_abi_version(void) { return FPM_AMD_ABI_VERSION; }

int fpm_pure_v3(const fpm_c32* O0, int64_t N_rows, int64_t N_cols, const fpm_c32* P0, int64_t n_rows, int64_t n_cols,
                const void* Y, int32_t y_dtype, int64_t L, const int32_t* led_idx, const float* A, double pratio,
                double epochs, double batch, const double opt[3], fpm_c32* O_out, fpm_c32* P_out,
                const fpm_opts* opts) {
    return run_guarded("fpm_pure_v3", [&] {
        const fpm_opts o = merged_opts(opts);
        if (!O0 || !P0 || !Y || !led_idx || !A || !opt || !O_out || !P_out) invalid("null array argument");
        (void)make_geom(N_rows, N_cols, n_rows, n_cols);  // shape errors before argument errors
        if (L < 1) invalid("L (number of images) must be >= 1");
        // doubles truncated to int as the mex does (cuFPM_pure_v3.cu:57-59)
        const int pr = (int)pratio, ne = (int)epochs, bs = (int)batch;
        if (pr < 1) invalid("pratio must be >= 1");
        if (ne < 0) invalid("numEpochs must be >= 0");
        if (bs < 1) invalid("batch must be >= 1");
        if (o.stitch != FPM_STITCH_ATOMIC && o.stitch != FPM_STITCH_OWNER) invalid("unknown stitch mode");
        if (o.lr_halve_from < 0) invalid("lr_halve_from must be >= 0");
        if (!(o.stop_loss >= 0.0)) invalid("stop_loss must be >= 0 (0: off)");
        const float lr = (float)opt[0];
        Geom g = make_geom(N_rows, N_cols, n_rows, n_cols);
        g.eng = engine_of(o.fft_engine);
        check_generic_options(g, g.eng, o.stitch);
        const float beta = (float)opt[1], eps = (float)opt[2];
        if (o.verbose) {
            std::printf("prhs4: %d, prhs5: %d, prhs6: %d \n", pr, ne, bs);
            std::printf("lr: %f, beta: %f, eps: %f \n", lr, beta, eps);
        }
        check_device();
        int dev = 0;
        FPM_HIP(hipGetDevice(&dev));
        const float2* tw = twiddle_table(dev);
        StreamHolder sh(o.stream);
        hipStream_t st = sh.s;
        const bool isdev = o.pointers_are_device != 0;
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        const uint32_t B = (uint32_t)std::min<int64_t>(bs, L);

        // working O and P (outputs double as the working arrays)
        DevBuf dO, dP;
        float2 *O, *P;
        if (isdev) {
            O = (float2*)O_out;
            P = (float2*)P_out;
            if ((const void*)O0 != (void*)O_out)
                FPM_HIP(hipMemcpyAsync(O, O0, NN * sizeof(float2), hipMemcpyDeviceToDevice, st));
            if ((const void*)P0 != (void*)P_out)
                FPM_HIP(hipMemcpyAsync(P, P0, nn * sizeof(float2), hipMemcpyDeviceToDevice, st));
        } else {
            dO.alloc(NN * sizeof(float2));
            dP.alloc(nn * sizeof(float2));
            O = dO.as<float2>();
            P = dP.as<float2>();
            FPM_HIP(hipMemcpyAsync(O, O0, NN * sizeof(float2), hipMemcpyHostToDevice, st));
            FPM_HIP(hipMemcpyAsync(P, P0, nn * sizeof(float2), hipMemcpyHostToDevice, st));
        }
        DevBuf hY, hLed, hA;
        const YStack y = prepare_y(Y, y_dtype, nn * (size_t)L, isdev, hY, st);
        const int2* led = (const int2*)upload(led_idx, 2 * (size_t)L, isdev, hLed, st);
        const float* dA = upload(A, nn, isdev, hA, st);

        int ran = 0;
        if (ne > 0) {
            Solve sv;
            sv.g = g;
            sv.y = y;
            sv.led = led;
            sv.A = dA;
            sv.O = O;
            sv.P = P;
            sv.L = (uint32_t)L;
            sv.B = B;
            sv.pr = pr;
            sv.stitch = o.stitch;
            sv.tw = tw;
            StageTimer tm(o.stage_times, st);
            const bool stop = o.stop_loss > 0.0;
            sv.init(ne, o.loss_per_epoch != nullptr || stop, lr, beta, eps, o.pupil_clamp, o.lr_halve_from, st, tm);
            const bool use_graph = o.graph != 0 && o.stage_times == nullptr;
            GraphCache graphs;
            int cur = 0;
            for (int e = 0; e < ne; ++e) {
                const bool last_epoch = e == ne - 1;
                if (use_graph) {
                    const int key = cur * 2 + (last_epoch ? 1 : 0);
                    hipGraphExec_t ex = graphs.get(key, st, [&] { sv.enqueue_epoch(cur, last_epoch, st, tm); });
                    FPM_HIP(hipGraphLaunch(ex, st));
                } else {
                    sv.enqueue_epoch(cur, last_epoch, st, tm);
                }
                cur = sv.parity_after(cur);
                ran = e + 1;
                if (stop && !last_epoch) {
                    // the epoch-end kernel has written this epoch's loss
                    double lv = 0.0;
                    FPM_HIP(hipMemcpyAsync(&lv, sv.ss.loss_hist + e, sizeof(double), hipMemcpyDeviceToHost, st));
                    FPM_HIP(hipStreamSynchronize(st));
                    if (lv <= o.stop_loss) break;
                }
            }
            if (o.loss_per_epoch) {
                FPM_HIP(hipMemcpyAsync(o.loss_per_epoch, sv.ss.loss_hist, (size_t)ran * sizeof(double),
                                       hipMemcpyDeviceToHost, st));
                for (int e = ran; e < ne; ++e) o.loss_per_epoch[e] = std::nan("");
            }
            FPM_HIP(hipStreamSynchronize(st));
            tm.flush();
        }
        if (o.epochs_run) *o.epochs_run = ran;
        if (!isdev) {
            FPM_HIP(hipMemcpyAsync(O_out, O, NN * sizeof(float2), hipMemcpyDeviceToHost, st));
            FPM_HIP(hipMemcpyAsync(P_out, P, nn * sizeof(float2), hipMemcpyDeviceToHost, st));
        }
        FPM_HIP(hipStreamSynchronize(st));
    });
}

int fpm_tiles_solve(const fpm_tile* tiles, int32_t ntiles, int64_t L, double pratio, double epochs, double batch,
                    const double opt[3], int32_t streams, const fpm_opts* opts, fpm_tiles_report* report) {
    return run_guarded("fpm_tiles_solve", [&] {
        const fpm_opts o = merged_opts(opts);
        if (!tiles || !opt) invalid("null argument");
        if (ntiles < 1 || ntiles > 4096) invalid("ntiles must be in [1, 4096]");
        if (L < 1) invalid("L (number of images) must be >= 1");
        const int pr = (int)pratio, ne = (int)epochs, bs = (int)batch;
        if (pr < 1) invalid("pratio must be >= 1");
        if (ne < 0) invalid("numEpochs must be >= 0");
        if (bs < 1) invalid("batch must be >= 1");
        if (streams < 0 || streams > 64) invalid("streams must be in [0, 64]");
        if (o.stitch != FPM_STITCH_ATOMIC && o.stitch != FPM_STITCH_OWNER) invalid("unknown stitch mode");
        if (o.lr_halve_from < 0) invalid("lr_halve_from must be >= 0");
        if (o.stop_loss != 0.0 || o.epochs_run) invalid("stop_loss / epochs_run are options of fpm_pure_v3 only");
        if (o.tile_launch != FPM_TILES_GROUPED && o.tile_launch != FPM_TILES_PER_TILE)
            invalid("tile_launch must be FPM_TILES_GROUPED or FPM_TILES_PER_TILE");
        const int eng = engine_of(o.fft_engine);
        std::vector<Geom> geo(ntiles);
        for (int t = 0; t < ntiles; ++t) {
            const fpm_tile& tl = tiles[t];
            if (!tl.O0 || !tl.O_out || !tl.P0 || !tl.P_out || !tl.Y || !tl.led_idx || !tl.A)
                invalid("tile " + std::to_string(t) + ": null array argument");
            if (tl.y_dtype != FPM_Y_F32_AMPLITUDE && tl.y_dtype != FPM_Y_U16_RAW)
                invalid("tile " + std::to_string(t) + ": y_dtype must be FPM_Y_F32_AMPLITUDE or FPM_Y_U16_RAW");
            geo[t] = make_geom(tl.N, tl.N, tl.n, tl.n);
            geo[t].eng = eng;
        }
        const float lr = (float)opt[0], beta = (float)opt[1], eps = (float)opt[2];
        check_device();
        int dev = 0;
        FPM_HIP(hipGetDevice(&dev));
        const float2* tw = twiddle_table(dev);
        StreamHolder sh(o.stream);
        hipStream_t st = sh.s;
        const uint32_t B = (uint32_t)std::min<int64_t>(bs, L);
        StageTimer none(nullptr, st);

        // launch sequences: groups of tiles of one size and image type (two or
        // more, groupable), and single tiles
        std::vector<std::vector<int>> groups;
        std::vector<int> singles;
        {
            std::map<std::pair<int64_t, int32_t>, std::vector<int>> by;
            for (int t = 0; t < ntiles; ++t) {
                if (o.tile_launch == FPM_TILES_GROUPED && GroupSolve::groupable(geo[t], B, o.stitch))
                    by[{tiles[t].n, tiles[t].y_dtype}].push_back(t);
                else
                    singles.push_back(t);
            }
            // a launch takes at most 65535 pages per grid dimension (T b of them)
            const size_t max_group = std::max<size_t>(1, 65535 / B);
            for (auto& kv : by) {
                const std::vector<int>& m = kv.second;
                for (size_t a = 0; a < m.size(); a += max_group) {
                    const size_t e = std::min(m.size(), a + max_group);
                    if (e - a >= 2) groups.emplace_back(m.begin() + a, m.begin() + e);
                    else singles.push_back(m[a]);
                }
            }
        }
        std::vector<std::unique_ptr<GroupSolve>> gsv;
        std::vector<std::unique_ptr<Solve>> sv;
        std::vector<std::unique_ptr<DevBuf>> ybuf;
        struct Seq {
            std::function<void(int, bool, hipStream_t)> enqueue;
            double work;
        };
        std::vector<Seq> seqs;
        auto tile_work = [&](int t) {
            const Geom& g = geo[t];
            return (double)g.Nr * g.Nc + (double)B * g.nr * g.nc;
        };
        int grouped_tiles = 0;
        for (const auto& members : groups) {
            gsv.push_back(std::make_unique<GroupSolve>());
            GroupSolve& gs = *gsv.back();
            gs.g = geo[members[0]];
            gs.L = (uint32_t)L;
            gs.B = B;
            gs.pr = pr;
            gs.tw = tw;
            double w = 0.0;
            for (int t : members) w += tile_work(t);
            if (ne > 0) {
                gs.init(tiles, members, ne, lr, beta, eps, o.pupil_clamp, o.lr_halve_from, st);
                seqs.push_back({[&gs](int cur, bool last, hipStream_t s) { gs.enqueue_epoch(cur, last, s); }, w});
            } else {
                for (int t : members) {
                    const fpm_tile& tl = tiles[t];
                    const Geom& g = geo[t];
                    if ((const void*)tl.O0 != (void*)tl.O_out)
                        FPM_HIP(hipMemcpyAsync(tl.O_out, tl.O0, (size_t)g.Nr * g.Nc * sizeof(float2),
                                               hipMemcpyDeviceToDevice, st));
                    if ((const void*)tl.P0 != (void*)tl.P_out)
                        FPM_HIP(hipMemcpyAsync(tl.P_out, tl.P0, (size_t)g.nr * g.nc * sizeof(float2),
                                               hipMemcpyDeviceToDevice, st));
                }
            }
            grouped_tiles += (int)members.size();
        }
        for (int t : singles) {
            const fpm_tile& tl = tiles[t];
            const Geom& g = geo[t];
            const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
            if ((const void*)tl.O0 != (void*)tl.O_out)
                FPM_HIP(hipMemcpyAsync(tl.O_out, tl.O0, NN * sizeof(float2), hipMemcpyDeviceToDevice, st));
            if ((const void*)tl.P0 != (void*)tl.P_out)
                FPM_HIP(hipMemcpyAsync(tl.P_out, tl.P0, nn * sizeof(float2), hipMemcpyDeviceToDevice, st));
            ybuf.push_back(std::make_unique<DevBuf>());
            sv.push_back(std::make_unique<Solve>());
            Solve& s = *sv.back();
            s.g = g;
            s.y = prepare_y(tl.Y, tl.y_dtype, nn * (size_t)L, true, *ybuf.back(), st);
            s.led = (const int2*)tl.led_idx;
            s.A = tl.A;
            s.O = (float2*)tl.O_out;
            s.P = (float2*)tl.P_out;
            s.L = (uint32_t)L;
            s.B = B;
            s.pr = pr;
            s.stitch = o.stitch;
            s.tw = tw;
            if (ne > 0) {
                s.init(ne, tl.loss_per_epoch != nullptr, lr, beta, eps, o.pupil_clamp, o.lr_halve_from, st, none);
                seqs.push_back(
                    {[&s, &none](int cur, bool last, hipStream_t x) { s.enqueue_epoch(cur, last, x, none); },
                     tile_work(t)});
            }
        }
        if (report) {
            std::memset(report, 0, sizeof(*report));
            report->sequences = (int32_t)(groups.size() + singles.size());
            report->grouped_tiles = grouped_tiles;
        }
        if (ne > 0) {
            const int nseq = (int)seqs.size();
            const uint32_t steps = (uint32_t)((L + B - 1) / B);
            // candidate stream counts (at most one per launch sequence);
            // sequences largest first, each to the least loaded stream
            std::vector<int> cands;
            if (streams > 0) {
                cands.push_back(std::min(streams, nseq));
            } else {
                for (int k : {1, 2, 4, 8})
                    if (k <= nseq && (cands.empty() || cands.back() != k)) cands.push_back(k);
                if (cands.back() < std::min(8, nseq)) cands.push_back(std::min(8, nseq));
            }
            const int kmax = *std::max_element(cands.begin(), cands.end());
            std::vector<int> order(nseq);
            for (int q = 0; q < nseq; ++q) order[q] = q;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return seqs[a].work > seqs[b].work; });
            struct Streams {
                std::vector<hipStream_t> s;
                std::vector<hipEvent_t> e;
                ~Streams() {
                    for (auto x : s) (void)hipStreamDestroy(x);
                    for (auto x : e) (void)hipEventDestroy(x);
                }
            } xs;
            for (int k = 1; k < kmax; ++k) {
                hipStream_t x;
                FPM_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
                xs.s.push_back(x);
            }
            for (int k = 0; k < kmax; ++k) {
                hipEvent_t e;
                FPM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                xs.e.push_back(e);
            }
            auto stream_of = [&](int k) { return k == 0 ? st : xs.s[k - 1]; };
            // all sequences' epoch on K streams: fork from st, join back into st
            auto enqueue_all = [&](int K, int cur, bool last) {
                std::vector<std::vector<int>> bucket(K);
                std::vector<double> load(K, 0.0);
                for (int q : order) {
                    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                    bucket[k].push_back(q);
                    load[k] += seqs[q].work;
                }
                if (K > 1) {
                    FPM_HIP(hipEventRecord(xs.e[0], st));
                    for (int k = 1; k < K; ++k) FPM_HIP(hipStreamWaitEvent(stream_of(k), xs.e[0], 0));
                }
                for (int k = 0; k < K; ++k)
                    for (int q : bucket[k]) seqs[q].enqueue(cur, last, stream_of(k));
                if (K > 1) {
                    for (int k = 1; k < K; ++k) {
                        FPM_HIP(hipEventRecord(xs.e[k], stream_of(k)));
                        FPM_HIP(hipStreamWaitEvent(st, xs.e[k], 0));
                    }
                }
            };
            hipEvent_t ta, tb;
            FPM_HIP(hipEventCreate(&ta));
            FPM_HIP(hipEventCreate(&tb));
            struct EvPair {
                hipEvent_t a, b;
                ~EvPair() {
                    (void)hipEventDestroy(a);
                    (void)hipEventDestroy(b);
                }
            } evp{ta, tb};
            GraphCache graphs;
            const bool tune = cands.size() > 1;
            std::vector<double> cand_ms(cands.size(), 0.0);
            int K = tune ? -1 : cands[0];
            int cur = 0;
            // With tuning, epoch 0 runs the sequences one after another on the
            // call's stream with an event after each: their measured times
            // replace the size model in the largest-first assignment of every
            // candidate stream count timed after it.
            const bool measure = tune && nseq > 1;
            const int skip = measure ? 1 : 0;
            for (int e = 0; e < ne; ++e) {
                const bool last = e == ne - 1;
                if (measure && e == 0) {
                    struct Evs {
                        std::vector<hipEvent_t> v;
                        ~Evs() {
                            for (auto x : v) (void)hipEventDestroy(x);
                        }
                    } ev;
                    for (int q = 0; q <= nseq; ++q) {
                        hipEvent_t x;
                        FPM_HIP(hipEventCreate(&x));
                        ev.v.push_back(x);
                    }
                    FPM_HIP(hipEventRecord(ev.v[0], st));
                    for (int i = 0; i < nseq; ++i) {
                        seqs[order[i]].enqueue(cur, last, st);
                        FPM_HIP(hipEventRecord(ev.v[i + 1], st));
                    }
                    FPM_HIP(hipEventSynchronize(ev.v[nseq]));
                    for (int i = 0; i < nseq; ++i) {
                        float ms = 0.f;
                        FPM_HIP(hipEventElapsedTime(&ms, ev.v[i], ev.v[i + 1]));
                        seqs[order[i]].work = (double)ms;
                    }
                    std::stable_sort(order.begin(), order.end(),
                                     [&](int a, int b) { return seqs[a].work > seqs[b].work; });
                    if (last) K = 1;
                    cur = (int)((cur + steps) & 1u);
                    continue;
                }
                const int ce = e - skip;  // candidate index
                const bool timing = K < 0 && ce < (int)cands.size();
                const int Ke = timing ? cands[ce] : K;
                const int key = (Ke * 2 + cur) * 2 + (last ? 1 : 0);
                hipGraphExec_t ex = graphs.get(key, st, [&] { enqueue_all(Ke, cur, last); });
                if (timing) {
                    FPM_HIP(hipEventRecord(ta, st));
                    FPM_HIP(hipGraphLaunch(ex, st));
                    FPM_HIP(hipEventRecord(tb, st));
                    FPM_HIP(hipEventSynchronize(tb));
                    float ms = 0.f;
                    FPM_HIP(hipEventElapsedTime(&ms, ta, tb));
                    cand_ms[ce] = ms;
                    if (ce == (int)cands.size() - 1 || last) {
                        const int nt = ce + 1;
                        K = cands[(int)(std::min_element(cand_ms.begin(), cand_ms.begin() + nt) - cand_ms.begin())];
                    }
                } else {
                    FPM_HIP(hipGraphLaunch(ex, st));
                }
                cur = (int)((cur + steps) & 1u);
            }
            for (size_t q = 0; q < groups.size(); ++q) gsv[q]->finish(tiles, groups[q], ne, st);
            for (size_t q = 0; q < singles.size(); ++q) {
                const fpm_tile& tl = tiles[singles[q]];
                if (tl.loss_per_epoch)
                    FPM_HIP(hipMemcpyAsync(tl.loss_per_epoch, sv[q]->ss.loss_hist, (size_t)ne * sizeof(double),
                                           hipMemcpyDeviceToHost, st));
            }
            if (report) {
                report->streams = K;
                report->tuned = tune ? (int32_t)std::min<size_t>(cands.size(), (size_t)std::max(0, ne - skip)) : 0;
                for (size_t k = 0; k < cands.size() && k < 4; ++k) {
                    report->cand[k] = cands[k];
                    report->cand_ms[k] = cand_ms[k];
                }
            }
            // the graphs, streams and events of this block are released on
            // leaving it: nothing of theirs may still run
            FPM_HIP(hipStreamSynchronize(st));
        }
        FPM_HIP(hipStreamSynchronize(st));
    });
}

int fpm_fused_gradient(const fpm_c32* O_in, int64_t N_rows, int64_t N_cols, const fpm_c32* P_in, int64_t n_rows,
                       int64_t n_cols, const void* Y, int32_t y_dtype, int64_t L, const int32_t* led_idx,
                       double pratio, fpm_c32* G_O, fpm_c32* G_P, double* loss_out, const fpm_opts* opts) {
    return run_guarded("fpm_fused_gradient", [&] {
        const fpm_opts o = merged_opts(opts);
        if (!O_in || !P_in || !Y || !led_idx || !G_O || !G_P) invalid("null array argument");
        Geom g = make_geom(N_rows, N_cols, n_rows, n_cols);
        g.eng = engine_of(o.fft_engine);
        if (o.stitch != FPM_STITCH_ATOMIC && o.stitch != FPM_STITCH_OWNER) invalid("unknown stitch mode");
        check_generic_options(g, g.eng, o.stitch);
        if (L < 1) invalid("L (number of images) must be >= 1");
        const int pr = (int)pratio;
        if (pr < 1) invalid("pratio must be >= 1");
        check_device();
        int dev = 0;
        FPM_HIP(hipGetDevice(&dev));
        const float2* tw = twiddle_table(dev);
        StreamHolder sh(o.stream);
        hipStream_t st = sh.s;
        const bool isdev = o.pointers_are_device != 0;
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        DevBuf hO, hP, hY, hLed, pcopy, S0, S1, gp, gout, lossd;
        const float2* O = (const float2*)upload((const fpm_c32*)O_in, NN, isdev, hO, st);
        // P is read-only here; the back-projection takes a mutable pointer
        pcopy.alloc(nn * sizeof(float2));
        FPM_HIP(hipMemcpyAsync(pcopy.p, P_in, nn * sizeof(float2), isdev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               st));
        const YStack y = prepare_y(Y, y_dtype, nn * (size_t)L, isdev, hY, st);
        const int2* led = (const int2*)upload(led_idx, 2 * (size_t)L, isdev, hLed, st);
        Workspace ws;
        ws.alloc(nn * L * sizeof(float2));
        S0.alloc(NN * sizeof(float2));
        S1.alloc(NN * sizeof(float2));
        float2* GPd;
        if (isdev) {
            GPd = (float2*)G_P;
        } else {
            gp.alloc(nn * sizeof(float2));
            GPd = gp.as<float2>();
        }
        if (loss_out) {
            lossd.alloc(sizeof(double));
            FPM_HIP(hipMemsetAsync(lossd.p, 0, sizeof(double), st));
        }
        StageTimer tm(o.stage_times, st);
        float2* GOd;
        if (isdev) {
            GOd = (float2*)G_O;
        } else {
            gout.alloc(NN * sizeof(float2));
            GOd = gout.as<float2>();
        }
        // the loss stencil's sign decisions (optional diagnostic output)
        DevBuf sgn;
        uint8_t* signs = nullptr;
        if (o.stencil_signs) {
            if (isdev) {
                signs = o.stencil_signs;
            } else {
                sgn.alloc(nn * (size_t)L);
                signs = sgn.as<uint8_t>();
            }
        }
        if (g.generic) {
            // generic sides (generic.hip): FP32 mixed-radix transforms, owner stitch
            DevBuf eb;
            if (!generic_fused_pages(g, signs != nullptr)) eb.alloc(nn * (size_t)L * sizeof(float));
            generic_prepare(g);
            float2* S = S0.as<float2>();
            float2* lat = ws.lat;
            float2* rec = ws.rec;
            const uint32_t b = (uint32_t)L;
            const bool fused = generic_fused_pages(g, signs != nullptr);
            fft2_generic(O, S, g.Nr, g.Nc, 1, false, st);
            generic_inverse_pages(S, pcopy.as<float2>(), led, 0, b, g, lat, st, nullptr, 1, fused);
            generic_stencil_pages(lat, y, 0, b, pr, g, eb.as<float>(), rec, loss_out ? lossd.as<double>() : nullptr, st,
                                  nullptr, 1, signs, fused);
            generic_forward_pages(rec, b, g, st, fused);
            generic_stitch(rec, pcopy.as<float2>(), led, 0, b, g, S1.as<float2>(), st);
            generic_pupil(rec, S, led, 0, b, g, pcopy.as<float2>(), nullptr, nullptr, RmsArgs{0.f, 0.f, 0.f, 0.f},
                          GPd, st);
            fft2_generic(S1.as<float2>(), GOd, g.Nr, g.Nc, 1, true, st);
        } else {
            tm.begin(FPM_STAGE_SPECTRUM_INIT);
            fft_contig(O, S0.as<float2>(), g.Nr, g.Nc, false, tw, st, g.eng);
            fft_cols_forward(S0.as<float2>(), g.Nr, g.Nc, 1, g.pN, tw, st, nullptr, nullptr, nullptr, g.eng);
            tm.end(1 + col_launches(g.pN));
            minibatch_backward(g, S0.as<float2>(), pcopy.as<float2>(), led, 0, (uint32_t)L, y, pr, ws, S1.as<float2>(),
                               o.stitch, nullptr, nullptr, RmsArgs{0.f, 0.f, 0.f, 0.f}, GPd,
                               loss_out ? lossd.as<double>() : nullptr, tw, st, tm, nullptr, nullptr, signs);
            fft_contig(S1.as<float2>(), GOd, g.Nr, g.Nc, true, tw, st, g.eng);
        }
        if (!isdev) {
            FPM_HIP(hipMemcpyAsync(G_O, GOd, NN * sizeof(float2), hipMemcpyDeviceToHost, st));
            FPM_HIP(hipMemcpyAsync(G_P, GPd, nn * sizeof(float2), hipMemcpyDeviceToHost, st));
        }
        if (loss_out) FPM_HIP(hipMemcpyAsync(loss_out, lossd.p, sizeof(double), hipMemcpyDeviceToHost, st));
        if (signs && !isdev)
            FPM_HIP(hipMemcpyAsync(o.stencil_signs, signs, nn * (size_t)L, hipMemcpyDeviceToHost, st));
        FPM_HIP(hipStreamSynchronize(st));
        tm.flush();
    });
}

// Scratch of fpm_fused_gradient_noalloc: the two object spectra, the two sets
// of per-LED pages and the pupil copy that fpm_fused_gradient allocates, in
// one static device array.  Each GPU the library runs on holds its own copy
// until the process ends; every call on a GPU shares that GPU's copy.
static __device__ __align__(256) unsigned char scratch[FPM_NOALLOC_SCRATCH_BYTES];

// Address of this device's copy of `scratch`: looked up once per device (with
// that device current) on the entry point's first call there.  In host code
// the array's name is only a stand-in; kernels and copies take this address.
unsigned char* noalloc_scratch(int device) {
    static std::mutex mu;
    static std::map<int, unsigned char*> ptrs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(device);
    if (it != ptrs.end()) return it->second;
    void* p = nullptr;
    FPM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) invalid("hipGetSymbolAddress returned no address for the scratch array", FPM_E_HIP);
    ptrs[device] = (unsigned char*)p;
    return (unsigned char*)p;
}

int fpm_fused_gradient_noalloc(const fpm_c32* O_in, int64_t N_rows, int64_t N_cols, const fpm_c32* P_in,
                               int64_t n_rows, int64_t n_cols, const float* Y, int64_t L, const int32_t* led_idx,
                               double pratio, fpm_c32* G_O, fpm_c32* G_P, double* loss_out, void* stream) {
    return run_guarded("fpm_fused_gradient_noalloc", [&] {
        if (!O_in || !P_in || !Y || !led_idx || !G_O || !G_P) invalid("null array argument");
        Geom g = make_geom(N_rows, N_cols, n_rows, n_cols);
        if (g.generic)
            invalid("fpm_fused_gradient_noalloc takes power-of-two sides only (the generic path allocates its "
                    "twiddle tables)", FPM_E_UNSUPPORTED);
        g.eng = ENG_F32;
        if (L < 1) invalid("L (number of images) must be >= 1");
        const int pr = (int)pratio;
        if (pr < 1) invalid("pratio must be >= 1");
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t spec = up(NN * sizeof(float2)), pages = up(nn * (size_t)L * sizeof(float2));
        const size_t need = 2 * spec + 2 * pages + up(nn * sizeof(float2));
        if (need > (size_t)FPM_NOALLOC_SCRATCH_BYTES)
            invalid("this call needs " + std::to_string(need) + " bytes of scratch (2 N_rows N_cols + 2 L n_rows n_cols + "
                    "n_rows n_cols complex values); the static scratch holds " +
                    std::to_string((size_t)FPM_NOALLOC_SCRATCH_BYTES));
        check_device();
        int dev = 0;
        FPM_HIP(hipGetDevice(&dev));
        unsigned char* base = noalloc_scratch(dev);
        const float2* tw = twiddle_table(dev);
        hipStream_t st = (hipStream_t)stream;
        float2* S0 = (float2*)base;
        float2* S1 = (float2*)(base + spec);
        Workspace ws;
        ws.lat = (float2*)(base + 2 * spec);
        ws.rec = (float2*)(base + 2 * spec + pages);
        float2* pcopy = (float2*)(base + 2 * spec + 2 * pages);
        // P is read-only here; the back-projection takes a mutable pointer
        FPM_HIP(hipMemcpyAsync(pcopy, P_in, nn * sizeof(float2), hipMemcpyDeviceToDevice, st));
        // the loss accumulates into the caller's slot: cleared by a fill kernel
        if (loss_out) launch_zero((float2*)loss_out, 1, st);
        YStack y{};
        y.data = Y;
        y.u16 = 0;
        StageTimer tm(nullptr, st);
        fft_contig((const float2*)O_in, S0, g.Nr, g.Nc, false, tw, st, g.eng);
        fft_cols_forward(S0, g.Nr, g.Nc, 1, g.pN, tw, st, nullptr, nullptr, nullptr, g.eng);
        minibatch_backward(g, S0, pcopy, (const int2*)led_idx, 0, (uint32_t)L, y, pr, ws, S1, FPM_STITCH_OWNER,
                           nullptr, nullptr, RmsArgs{0.f, 0.f, 0.f, 0.f}, (float2*)G_P, loss_out, tw, st, tm, nullptr,
                           nullptr);
        fft_contig(S1, (float2*)G_O, g.Nr, g.Nc, true, tw, st, g.eng);
    });
}

int fpm_forward_intensity(const fpm_c32* O_in, int64_t N_rows, int64_t N_cols, const fpm_c32* P_in, int64_t n_rows,
                          int64_t n_cols, int64_t L, const int32_t* led_idx, double pratio, int32_t squared,
                          float* out, int32_t pointers_are_device, void* stream) {
    return run_guarded("fpm_forward_intensity", [&] {
        if (!O_in || !P_in || !led_idx || !out) invalid("null array argument");
        const Geom g = make_geom(N_rows, N_cols, n_rows, n_cols);
        if (L < 1) invalid("L (number of images) must be >= 1");
        const int pr = (int)pratio;
        if (pr < 1) invalid("pratio must be >= 1");
        check_device();
        int dev = 0;
        FPM_HIP(hipGetDevice(&dev));
        const float2* tw = twiddle_table(dev);
        StreamHolder sh(stream);
        hipStream_t st = sh.s;
        const bool isdev = pointers_are_device != 0;
        const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
        DevBuf hO, hP, hLed, S0, lat, outd;
        const float2* O = (const float2*)upload(O_in, NN, isdev, hO, st);
        const float2* P = (const float2*)upload(P_in, nn, isdev, hP, st);
        const int2* led = (const int2*)upload(led_idx, 2 * (size_t)L, isdev, hLed, st);
        S0.alloc(NN * sizeof(float2));
        const uint32_t chunk = (uint32_t)std::min<int64_t>(L, 64);
        lat.alloc(nn * chunk * sizeof(float2));
        float* od = out;
        if (!isdev) {
            outd.alloc(nn * (size_t)L * sizeof(float));
            od = outd.as<float>();
        }
        const float scale = (float)((double)nn * (double)pr * (double)pr);
        if (g.generic) {
            generic_prepare(g);
            fft2_generic(O, S0.as<float2>(), g.Nr, g.Nc, 1, false, st);
            for (int64_t l0 = 0; l0 < L; l0 += chunk) {
                const uint32_t b = (uint32_t)std::min<int64_t>(chunk, L - l0);
                generic_crop(S0.as<float2>(), P, led, (uint32_t)l0, b, g, lat.as<float2>(), st);
                fft2_generic(lat.as<float2>(), lat.as<float2>(), g.nr, g.nc, b, true, st);
                launch_abs_scaled(lat.as<float2>(), nn * b, scale, squared, od + nn * (size_t)l0, st);
            }
        } else {
            fft_contig(O, S0.as<float2>(), g.Nr, g.Nc, false, tw, st);
            fft_cols_forward(S0.as<float2>(), g.Nr, g.Nc, 1, g.pN, tw, st, nullptr);
            for (int64_t l0 = 0; l0 < L; l0 += chunk) {
                const uint32_t b = (uint32_t)std::min<int64_t>(chunk, L - l0);
                const CropArgs ca{S0.as<float2>(), P, led, (uint32_t)l0, g.nr, g.nc, g.Nr, g.Nc, g.pN};
                fft_cols_inverse(lat.as<float2>(), g.nr, g.nc, b, g.pn, tw, st, nullptr, &ca);
                fft_contig(lat.as<float2>(), lat.as<float2>(), g.nr, g.nc * b, true, tw, st);
                launch_abs_scaled(lat.as<float2>(), nn * b, scale, squared, od + nn * (size_t)l0, st);
            }
        }
        if (!isdev) FPM_HIP(hipMemcpyAsync(out, od, nn * (size_t)L * sizeof(float), hipMemcpyDeviceToHost, st));
        FPM_HIP(hipStreamSynchronize(st));
    });
}

int fpm_fft2(const fpm_c32* in, fpm_c32* out, int64_t rows, int64_t cols, int64_t batch, int32_t direction,
             int32_t pointers_are_device, void* stream) {
    return fpm_fft2_engine(in, out, rows, cols, batch, direction, FPM_FFT_F32, pointers_are_device, stream);
}

int fpm_fft2_engine(const fpm_c32* in, fpm_c32* out, int64_t rows, int64_t cols, int64_t batch, int32_t direction,
                    int32_t engine, int32_t pointers_are_device, void* stream) {
    return run_guarded("fpm_fft2", [&] {
        const int eng = engine_of(engine);
        if (!in || !out) invalid("null array argument");
        check_side(rows, "rows");
        check_side(cols, "cols");
        if (batch < 1 || batch > 65535) invalid("batch must be in [1, 65535]");
        if (direction != -1 && direction != 1) invalid("direction must be -1 (forward) or +1 (inverse)");
        if (eng == ENG_SPLIT16 && !(is_pow2(rows) && is_pow2(cols)))
            invalid("the split-FP16 engine needs power-of-two sides (the generic transform is FP32)",
                    FPM_E_UNSUPPORTED);
        check_device();
        int dev = 0;
        FPM_HIP(hipGetDevice(&dev));
        const float2* tw = twiddle_table(dev);
        StreamHolder sh(stream);
        hipStream_t st = sh.s;
        const bool isdev = pointers_are_device != 0;
        const size_t total = (size_t)rows * cols * batch;
        const ColPerm p = make_colperm((uint32_t)cols);
        DevBuf din, dout, tmp;
        const float2* src = (const float2*)upload(in, total, isdev, din, st);
        float2* dst;
        if (isdev) {
            dst = (float2*)out;
        } else {
            dout.alloc(total * sizeof(float2));
            dst = dout.as<float2>();
        }
        const uint32_t R = (uint32_t)rows, C = (uint32_t)cols, Bt = (uint32_t)batch;
        if (!is_pow2(rows) || !is_pow2(cols)) {
            // a side that is not a power of two: the generic mixed-radix transform (FP32)
            if (!generic_side(rows) || !generic_side(cols))
                invalid("with a side that is not a power of two both sides must be even and in [16, 16384]",
                        FPM_E_UNSUPPORTED);
            fft2_generic(src, dst, R, C, Bt, direction > 0, st);
        } else if (direction < 0) {
            fft_contig(src, dst, R, C * Bt, false, tw, st, eng);
            if (p.M2 == 1) {
                fft_cols_forward(dst, R, C, Bt, p, tw, st, nullptr, nullptr, nullptr, eng);
            } else {
                tmp.alloc(total * sizeof(float2));
                fft_cols_forward(dst, R, C, Bt, p, tw, st, tmp.as<float2>(), nullptr, nullptr, eng);
                FPM_HIP(hipMemcpyAsync(dst, tmp.p, total * sizeof(float2), hipMemcpyDeviceToDevice, st));
            }
        } else {
            if (p.M2 == 1) {
                if (src != dst) FPM_HIP(hipMemcpyAsync(dst, src, total * sizeof(float2), hipMemcpyDeviceToDevice, st));
                fft_cols_inverse(dst, R, C, Bt, p, tw, st, nullptr, nullptr, nullptr, nullptr, nullptr, eng);
            } else {
                const float2* nat = src;
                if (src == dst) {
                    tmp.alloc(total * sizeof(float2));
                    FPM_HIP(hipMemcpyAsync(tmp.p, src, total * sizeof(float2), hipMemcpyDeviceToDevice, st));
                    nat = tmp.as<float2>();
                }
                fft_cols_inverse(dst, R, C, Bt, p, tw, st, nat, nullptr, nullptr, nullptr, nullptr, eng);
            }
            fft_contig(dst, dst, R, C * Bt, true, tw, st, eng);
        }
        if (!isdev) FPM_HIP(hipMemcpyAsync(out, dst, total * sizeof(float2), hipMemcpyDeviceToHost, st));
        FPM_HIP(hipStreamSynchronize(st));
    });
}

int fpm_prepare_amplitude(const uint16_t* raw, int64_t count, float* amp_out, int32_t pointers_are_device,
                          void* stream, uint16_t* min_out, uint16_t* max_out) {
    return run_guarded("fpm_prepare_amplitude", [&] {
        if (!raw || !amp_out) invalid("null array argument");
        if (count < 1) invalid("count must be >= 1");
        check_device();
        StreamHolder sh(stream);
        hipStream_t st = sh.s;
        const bool isdev = pointers_are_device != 0;
        DevBuf hraw, dout;
        const YStack y = prepare_y(raw, FPM_Y_U16_RAW, (size_t)count, isdev, hraw, st);
        float* od = amp_out;
        if (!isdev) {
            dout.alloc((size_t)count * sizeof(float));
            od = dout.as<float>();
        }
        launch_u16_to_amp((const uint16_t*)y.data, (size_t)count, y.mn, y.range, od, st);
        if (!isdev) FPM_HIP(hipMemcpyAsync(amp_out, od, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, st));
        FPM_HIP(hipStreamSynchronize(st));
        if (min_out) *min_out = (uint16_t)y.mn;
        if (max_out) *max_out = (uint16_t)(y.mn + (uint32_t)y.range);
    });
}

}  // extern "C"
