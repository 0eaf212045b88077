// cnn_api.hip -- the warm-start network: a U-Net of 3x3 convolutions that
// maps low-resolution image features to the initial amplitude and phase of
// the object, trained and run by the kernels of cnn_kernels.hip.  C-ABI
// entry points fpm_cnn_* (include/fpm_amd.h); no counterpart in the
// reference, whose solve starts from an upsampled centre image.
//
// Architecture (levels L, base channels C, upsampling u):
//   encoder level l = 0..L-1 at resolution n / 2^l, C 2^l channels:
//       [pool 2x2 unless l = 0] -> conv+ReLU -> conv+ReLU          (e_l)
//   decoder level l = L-2..0:
//       upsample 2x -> conv+ReLU (d_l,a) ; s_l = d_l,a + e_l ; conv+ReLU (d_l,b)
//   head: conv (C -> 2 u^2) -> depth-to-space u -> amplitude, phase at u n.
//   pupil head (K = pupil_terms > 0): mean of e_{L-1} over its pixels ->
//       64 hidden (ReLU) -> K coefficients of the pupil's phase
// Parameters (FP32, one vector): per conv in the order above, W (Cout x 9 x
// Cin, Cin of the first conv rounded up to a multiple of 32) then b (Cout);
// then, with a pupil head, W1 (64 x Cb), b1 (64), W2 (K x 64), b2 (K), where
// Cb = C 2^(L-1) is the bottom level's channel count.
#include "../../include/fpm_amd.h"
#include "cnn_kernels.h"

#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <string>
#include <vector>

namespace fpm {
void set_error(const std::string& msg);
}

namespace fpm {
namespace cnn {
namespace {

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = b ? b : 16;
        FPM_HIP(hipMalloc(&p, bytes));
    }
    template <class T> T* as() const { return (T*)p; }
};

struct Conv {
    int cin, cout, relu;
    int64_t woff, boff;  // offsets into the parameter vector
    int H, W;            // resolution it runs at
};

struct Tensor {  // BF16 activation, B x H x W x C
    int H, W, C;
    Buf act;   // forward value
    Buf grad;  // FP32 gradient
};

struct Net {
    fpm_cnn_desc d{};
    int cin_pad = 32;
    std::vector<Conv> convs;
    int64_t nparam = 0;
    Buf params, m, v, grads, wbf;  // FP32 params, Adam moments, gradients; BF16 copy
    int64_t step = 0;
    // activations for max_batch
    int B = 0;
    // tensor ids: input, per encoder level pool-out (l>0), e_a, e_b; per decoder
    // level up, d_a, s, d_b; head output (FP32)
    std::vector<std::unique_ptr<Tensor>> T;
    int t_in = -1;
    std::vector<int> t_pool, t_ea, t_eb, t_up, t_da, t_s, t_db;
    Buf head_f, head_g;  // FP32 head output and its gradient
    Buf xcol, xt, gz, gzt, wt, lossd;  // lossd: image loss sum, pupil loss
    int K = 0;                     // pupil terms (0: no pupil head)
    int cb = 0;                    // channels of the bottom level
    int64_t poff = -1;             // offset of the pupil head's parameters
    Buf pws, pgpool;               // pupil head workspace and pooled gradient
    hipStream_t st = nullptr;

    int add_tensor(int H, int W, int C) {
        auto t = std::make_unique<Tensor>();
        t->H = H;
        t->W = W;
        t->C = C;
        T.push_back(std::move(t));
        return (int)T.size() - 1;
    }
    int add_conv(int cin, int cout, int relu, int H, int W) {
        Conv c{cin, cout, relu, nparam, 0, H, W};
        nparam += (int64_t)cout * 9 * cin;
        c.boff = nparam;
        nparam += cout;
        convs.push_back(c);
        return (int)convs.size() - 1;
    }
};

void check_desc(const fpm_cnn_desc& d) {
    auto bad = [](const std::string& w) { throw HipError{hipErrorInvalidValue, "fpm_cnn: " + w}; };
    if (d.levels < 1 || d.levels > 6) bad("levels must be in [1, 6]");
    if (d.base_channels < 32 || d.base_channels % 32 || (d.base_channels << (d.levels - 1)) > 256)
        bad("base_channels must be a multiple of 32 with base * 2^(levels-1) <= 256");
    if (d.in_channels < 1 || d.in_channels > 256) bad("in_channels must be in [1, 256]");
    if (d.n < 8 || d.n % (1 << (d.levels - 1)) || ((d.n >> (d.levels - 1)) * (d.n >> (d.levels - 1))) % 8)
        bad("n must be divisible by 2^(levels-1) with at least 8 pixels at the coarsest level");
    if (d.up < 1 || 2 * d.up * d.up % 32) bad("2 up^2 must be a multiple of 32 (up = 4, 8, ...)");
    if (d.max_batch < 1) bad("max_batch must be >= 1");
    if (d.pupil_terms < 0 || d.pupil_terms > kPupilMaxTerms) bad("pupil_terms must be in [0, 16]");
    if (d.pupil_terms > 0 && d.max_batch > kPupilMaxBatch) bad("max_batch must be <= 64 with a pupil head");
    if (!(d.pupil_weight >= 0.0f) || !std::isfinite(d.pupil_weight)) bad("pupil_weight must be finite and >= 0");
}

// builds the layer list (convs in parameter order) and the tensors
void build(Net& net) {
    const fpm_cnn_desc& d = net.d;
    const int L = d.levels, C = d.base_channels, n = d.n;
    net.cin_pad = (d.in_channels + 31) / 32 * 32;
    net.t_in = net.add_tensor(n, n, net.cin_pad);
    int prev_c = net.cin_pad;
    for (int l = 0; l < L; ++l) {
        const int h = n >> l, c = C << l;
        net.t_pool.push_back(l > 0 ? net.add_tensor(h, h, prev_c) : -1);
        net.add_conv(prev_c, c, 1, h, h);
        net.t_ea.push_back(net.add_tensor(h, h, c));
        net.add_conv(c, c, 1, h, h);
        net.t_eb.push_back(net.add_tensor(h, h, c));
        prev_c = c;
    }
    net.t_up.assign(L, -1);
    net.t_da.assign(L, -1);
    net.t_s.assign(L, -1);
    net.t_db.assign(L, -1);
    for (int l = L - 2; l >= 0; --l) {
        const int h = n >> l, c = C << l;
        net.t_up[l] = net.add_tensor(h, h, prev_c);
        net.add_conv(prev_c, c, 1, h, h);
        net.t_da[l] = net.add_tensor(h, h, c);
        net.t_s[l] = net.add_tensor(h, h, c);
        net.add_conv(c, c, 1, h, h);
        net.t_db[l] = net.add_tensor(h, h, c);
        prev_c = c;
    }
    net.add_conv(C, 2 * d.up * d.up, 0, n, n);  // head
    net.K = d.pupil_terms;
    net.cb = C << (L - 1);
    if (net.K > 0) {
        net.poff = net.nparam;
        net.nparam += (int64_t)kPupilHidden * net.cb + kPupilHidden + (int64_t)net.K * kPupilHidden + net.K;
    }
}

// pupil head parameters (FP32 master copy) or their gradients
struct PupilParams {
    float *W1, *b1, *W2, *b2;
};
PupilParams pupil_params(const Net& net, float* base) {
    float* p = base + net.poff;
    PupilParams q;
    q.W1 = p;
    q.b1 = q.W1 + (size_t)kPupilHidden * net.cb;
    q.W2 = q.b1 + kPupilHidden;
    q.b2 = q.W2 + (size_t)net.K * kPupilHidden;
    return q;
}

// conv index helpers (parameter order: 2 per encoder level, 2 per decoder level, head)
int conv_enc(const Net& net, int l, int k) { (void)net; return 2 * l + k; }
int conv_dec(const Net& net, int l, int k) { return 2 * net.d.levels + 2 * (net.d.levels - 2 - l) + k; }
int conv_head(const Net& net) { return (int)net.convs.size() - 1; }

void alloc_batch(Net& net, int B) {
    if (B <= net.B) return;
    net.B = B;
    size_t max_xcol = 0, max_xt = 0, max_gz = 0, max_wt = 0;
    for (auto& t : net.T) {
        const size_t cnt = (size_t)B * t->H * t->W * t->C;
        t->act.alloc(cnt * 2);
        t->grad.alloc(cnt * 4);
    }
    for (auto& c : net.convs) {
        const size_t Q = (size_t)B * c.H * c.W;
        // the im2col matrix only for layers the implicit weight gradient does not cover
        if (wgrad3x3_supported(c.W, c.cin, c.cout)) max_xt = std::max(max_xt, (size_t)c.cin * Q);
        else max_xcol = std::max(max_xcol, (size_t)9 * c.cin * Q);
        max_gz = std::max(max_gz, (size_t)c.cout * Q);
        max_wt = std::max(max_wt, (size_t)9 * c.cin * c.cout);
    }
    const size_t nh = (size_t)B * net.d.n * net.d.n * 2 * net.d.up * net.d.up;
    net.head_f.alloc(nh * 4);
    net.head_g.alloc(nh * 4);
    net.xcol.alloc(max_xcol * 2);
    net.xt.alloc(max_xt * 2);
    net.gz.alloc(max_gz * 2);
    net.gzt.alloc(max_gz * 2);
    net.wt.alloc(max_wt * 2);
    if (net.K > 0) {
        net.pws.alloc((size_t)B * (net.cb + kPupilHidden + net.K) * 4);
        net.pgpool.alloc((size_t)B * net.cb * 4);
    }
}

const float* bias_of(const Net& net, const Conv& c) { return net.params.as<float>() + c.boff; }
const bf16_t* w16_of(const Net& net, const Conv& c) { return net.wbf.as<bf16_t>() + c.woff; }

void run_conv(Net& net, int ci, const bf16_t* x, bf16_t* y, float* yf, int B) {
    const Conv& c = net.convs[ci];
    ConvArgs a{x, w16_of(net, c), bias_of(net, c), y, yf, B, c.H, c.W, c.cin, c.cout, c.relu};
    conv3x3(a, net.st);
}

bf16_t* act(Net& net, int t) { return net.T[t]->act.as<bf16_t>(); }
float* grad(Net& net, int t) { return net.T[t]->grad.as<float>(); }

// forward of a batch already in the input tensor; head output in head_f,
// pupil coefficients (with a pupil head) in the workspace and coef_out (may
// be null)
void forward(Net& net, int B, float* coef_out = nullptr) {
    const int L = net.d.levels;
    hipStream_t st = net.st;
    const bf16_t* x = act(net, net.t_in);
    for (int l = 0; l < L; ++l) {
        const int h = net.d.n >> l;
        if (l > 0) {
            avgpool2_fwd(act(net, net.t_eb[l - 1]), act(net, net.t_pool[l]), B, 2 * h, 2 * h,
                         net.T[net.t_eb[l - 1]]->C, st);
            x = act(net, net.t_pool[l]);
        }
        run_conv(net, conv_enc(net, l, 0), x, act(net, net.t_ea[l]), nullptr, B);
        run_conv(net, conv_enc(net, l, 1), act(net, net.t_ea[l]), act(net, net.t_eb[l]), nullptr, B);
        x = act(net, net.t_eb[l]);
    }
    if (net.K > 0) {
        const int hb = net.d.n >> (L - 1);
        const PupilParams q = pupil_params(net, net.params.as<float>());
        pupil_head_fwd(act(net, net.t_eb[L - 1]), B, hb * hb, net.cb, q.W1, q.b1, q.W2, q.b2, net.K,
                       net.pws.as<float>(), coef_out, st);
    }
    int prev = net.t_eb[L - 1];
    for (int l = L - 2; l >= 0; --l) {
        const int h = net.d.n >> l;
        upsample2_fwd(act(net, prev), act(net, net.t_up[l]), B, h / 2, h / 2, net.T[prev]->C, st);
        run_conv(net, conv_dec(net, l, 0), act(net, net.t_up[l]), act(net, net.t_da[l]), nullptr, B);
        const Tensor& s = *net.T[net.t_s[l]];
        add_bf16(act(net, net.t_da[l]), act(net, net.t_eb[l]), act(net, net.t_s[l]), (int64_t)B * s.H * s.W * s.C,
                 st);
        run_conv(net, conv_dec(net, l, 1), act(net, net.t_s[l]), act(net, net.t_db[l]), nullptr, B);
        prev = net.t_db[l];
    }
    run_conv(net, conv_head(net), act(net, prev), nullptr, net.head_f.as<float>(), B);
}

// backward of one conv: gy (FP32 gradient of its output y), x its input;
// accumulates the weight / bias gradients, writes the input gradient (FP32)
// to gx unless gx is null
void conv_backward(Net& net, int ci, const bf16_t* x, const bf16_t* y, const float* gy, float* gx, int B) {
    const Conv& c = net.convs[ci];
    hipStream_t st = net.st;
    const int64_t Q = (int64_t)B * c.H * c.W;
    float* gW = net.grads.as<float>() + c.woff;
    float* gb = net.grads.as<float>() + c.boff;
    bf16_t* gz = net.gz.as<bf16_t>();
    bf16_t* gzt = net.gzt.as<bf16_t>();
    grad_prep(gy, y, c.relu, gz, gzt, gb, Q, c.cout, st);
    if (wgrad3x3_supported(c.W, c.cin, c.cout)) {
        to_channel_major(x, net.xt.as<bf16_t>(), Q, c.cin, st);
        wgrad3x3(gzt, net.xt.as<bf16_t>(), gW, B, c.H, c.W, c.cin, c.cout, st);
    } else {
        im2col_t(x, net.xcol.as<bf16_t>(), B, c.H, c.W, c.cin, st);
        const int64_t ks = std::max<int64_t>(1024, (Q / 64 + 31) / 32 * 32);
        gemm_nt_atomic(gzt, net.xcol.as<bf16_t>(), gW, c.cout, 9 * c.cin, Q, ks, st);
    }
    if (gx) {
        flip_transpose_w(w16_of(net, c), net.wt.as<bf16_t>(), c.cout, c.cin, st);
        ConvArgs a{gz, net.wt.as<bf16_t>(), nullptr, nullptr, gx, B, c.H, c.W, c.cout, c.cin, 0};
        conv3x3(a, st);
    }
}

// backward through the whole net from head_g (and, with a pupil head, from
// the coefficient targets); parameter gradients in grads
void backward(Net& net, int B, const float* pupil_targets) {
    const int L = net.d.levels;
    hipStream_t st = net.st;
    fill_f32(net.grads.as<float>(), net.nparam, 0.0f, st);
    int last = L >= 2 ? net.t_db[0] : net.t_eb[0];
    conv_backward(net, conv_head(net), act(net, last), nullptr, net.head_g.as<float>(), grad(net, last), B);
    for (int l = 0; l <= L - 2; ++l) {
        const Tensor& s = *net.T[net.t_s[l]];
        const int64_t cnt = (int64_t)B * s.H * s.W * s.C;
        conv_backward(net, conv_dec(net, l, 1), act(net, net.t_s[l]), act(net, net.t_db[l]), grad(net, net.t_db[l]),
                      grad(net, net.t_s[l]), B);
        // s = d_a + e_b: both receive g_s (the encoder's skip gradient starts here)
        add_f32(grad(net, net.t_s[l]), grad(net, net.t_eb[l]), cnt, 0, st);
        conv_backward(net, conv_dec(net, l, 0), act(net, net.t_up[l]), act(net, net.t_da[l]), grad(net, net.t_s[l]),
                      grad(net, net.t_up[l]), B);
        const int prev = l == L - 2 ? net.t_eb[L - 1] : net.t_db[l + 1];
        const int h = net.d.n >> l;
        upsample2_bwd(grad(net, net.t_up[l]), grad(net, prev), B, h / 2, h / 2, net.T[prev]->C, st);
    }
    if (net.K > 0) {
        // the bottom level's output also feeds the pupil head
        const int hb = net.d.n >> (L - 1);
        const PupilParams q = pupil_params(net, net.params.as<float>());
        const PupilParams g = pupil_params(net, net.grads.as<float>());
        pupil_head_bwd(net.pws.as<float>(), pupil_targets, B, hb * hb, net.cb, q.W1, q.W2, net.K, net.d.pupil_weight,
                       g.W1, g.b1, g.W2, g.b2, net.pgpool.as<float>(), grad(net, net.t_eb[L - 1]),
                       net.lossd.as<double>() + 1, st);
    }
    for (int l = L - 1; l >= 0; --l) {
        const bf16_t* xin = l > 0 ? act(net, net.t_pool[l]) : act(net, net.t_in);
        conv_backward(net, conv_enc(net, l, 1), act(net, net.t_ea[l]), act(net, net.t_eb[l]), grad(net, net.t_eb[l]),
                      grad(net, net.t_ea[l]), B);
        conv_backward(net, conv_enc(net, l, 0), xin, act(net, net.t_ea[l]), grad(net, net.t_ea[l]),
                      l > 0 ? grad(net, net.t_pool[l]) : nullptr, B);
        if (l > 0) {
            // e_{l-1} feeds the pool and (below the bottom level) the skip
            const int h = net.d.n >> (l - 1);
            avgpool2_bwd(grad(net, net.t_pool[l]), grad(net, net.t_eb[l - 1]), B, h, h, net.T[net.t_eb[l - 1]]->C,
                         l - 1 <= L - 2 ? 1 : 0, st);
        }
    }
}

void load_features(Net& net, const float* feats, int B) {
    features_to_nhwc(feats, act(net, net.t_in), B, net.d.in_channels, net.d.n, net.d.n, net.cin_pad, net.st);
}

double loss_and_grad(Net& net, const float* feats, const float* targets, const float* pupil_targets, int B) {
    load_features(net, feats, B);
    forward(net, B);
    FPM_HIP(hipMemsetAsync(net.lossd.p, 0, 2 * sizeof(double), net.st));
    const int64_t count = (int64_t)B * 2 * (int64_t)net.d.up * net.d.n * net.d.up * net.d.n;
    head_mse(net.head_f.as<float>(), targets, net.head_g.as<float>(), net.lossd.as<double>(), B, net.d.n, net.d.n,
             net.d.up, 1.0 / (double)count, net.st);
    backward(net, B, pupil_targets);
    double loss[2] = {0.0, 0.0};
    FPM_HIP(hipMemcpyAsync(loss, net.lossd.p, 2 * sizeof(double), hipMemcpyDeviceToHost, net.st));
    FPM_HIP(hipStreamSynchronize(net.st));
    return loss[0] / (double)count + loss[1];
}

void check_call(const Net* net, const float* feats, const float* targets, const float* pupil_targets, int64_t batch) {
    if (!net || !feats || !targets) throw HipError{hipErrorInvalidValue, "null argument"};
    if (batch < 1 || batch > net->d.max_batch) throw HipError{hipErrorInvalidValue, "batch outside [1, max_batch]"};
    if ((net->K > 0) != (pupil_targets != nullptr))
        throw HipError{hipErrorInvalidValue, net->K > 0 ? "the network has a pupil head: pupil_targets required"
                                                        : "pupil_targets given to a network without a pupil head"};
}

int guarded(const char* fn, const std::function<void()>& body) {
    set_error("");
    try {
        body();
        return FPM_OK;
    } catch (const HipError& e) {
        set_error(std::string(fn) + ": " + e.what);
        if (e.code == hipErrorInvalidValue) return FPM_E_INVALID;
        return e.code == hipErrorOutOfMemory ? FPM_E_NOMEM : FPM_E_HIP;
    } catch (const std::exception& e) {
        set_error(std::string(fn) + ": " + e.what());
        return FPM_E_HIP;
    }
}

}  // namespace
}  // namespace cnn
}  // namespace fpm

using namespace fpm::cnn;

extern "C" {

int fpm_cnn_create(const fpm_cnn_desc* desc, uint64_t seed, fpm_cnn** out) {
    return guarded("fpm_cnn_create", [&] {
        if (!desc || !out) throw fpm::HipError{hipErrorInvalidValue, "null argument"};
        *out = nullptr;
        check_desc(*desc);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            (void)hipGetLastError();
            throw fpm::HipError{hipErrorNoDevice, "no HIP device available"};
        }
        auto net = std::make_unique<Net>();
        net->d = *desc;
        build(*net);
        FPM_HIP(hipStreamCreateWithFlags(&net->st, hipStreamNonBlocking));
        net->params.alloc(net->nparam * 4);
        net->m.alloc(net->nparam * 4);
        net->v.alloc(net->nparam * 4);
        net->grads.alloc(net->nparam * 4);
        net->wbf.alloc(net->nparam * 2);
        net->lossd.alloc(2 * sizeof(double));
        // He-normal weights (std sqrt(2 / fan_in) over the real input
        // channels), zero biases, from a seeded host generator
        std::vector<float> h(net->nparam, 0.0f);
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.0f, 1.0f);
        for (size_t k = 0; k < net->convs.size(); ++k) {
            const Conv& c = net->convs[k];
            const int real_cin = k == 0 ? desc->in_channels : c.cin;
            // the head starts at zero: the untrained network reproduces the plain
            // init exactly, and the first updates need not silence random output
            // (which they do by driving ReLUs dead)
            const float sd = c.relu ? std::sqrt(2.0f / (9.0f * (float)real_cin)) : 0.0f;
            for (int co = 0; co < c.cout; ++co)
                for (int tap = 0; tap < 9; ++tap)
                    for (int ci = 0; ci < c.cin; ++ci)
                        h[c.woff + ((int64_t)co * 9 + tap) * c.cin + ci] = ci < real_cin ? sd * nd(rng) : 0.0f;
        }
        if (net->K > 0) {
            // pupil head: He-normal hidden layer; zero output layer (the
            // untrained head predicts the ideal pupil, the plain init)
            const float sd = std::sqrt(2.0f / (float)net->cb);
            for (int64_t i = 0; i < (int64_t)kPupilHidden * net->cb; ++i) h[net->poff + i] = sd * nd(rng);
        }
        FPM_HIP(hipMemcpy(net->params.p, h.data(), net->nparam * 4, hipMemcpyHostToDevice));
        fill_f32(net->m.as<float>(), net->nparam, 0.0f, net->st);
        fill_f32(net->v.as<float>(), net->nparam, 0.0f, net->st);
        to_bf16(net->params.as<float>(), net->wbf.as<bf16_t>(), net->nparam, net->st);
        alloc_batch(*net, desc->max_batch);
        FPM_HIP(hipStreamSynchronize(net->st));
        *out = (fpm_cnn*)net.release();
    });
}

void fpm_cnn_destroy(fpm_cnn* h) {
    Net* net = (Net*)h;
    if (!net) return;
    if (net->st) {
        (void)hipStreamSynchronize(net->st);
        (void)hipStreamDestroy(net->st);
    }
    delete net;
}

int64_t fpm_cnn_num_params(const fpm_cnn* h) { return h ? ((const Net*)h)->nparam : -1; }

int fpm_cnn_get_params(fpm_cnn* h, float* out) {
    return guarded("fpm_cnn_get_params", [&] {
        Net* net = (Net*)h;
        if (!net || !out) throw fpm::HipError{hipErrorInvalidValue, "null argument"};
        FPM_HIP(hipMemcpyAsync(out, net->params.p, net->nparam * 4, hipMemcpyDeviceToHost, net->st));
        FPM_HIP(hipStreamSynchronize(net->st));
    });
}

int fpm_cnn_set_params(fpm_cnn* h, const float* in) {
    return guarded("fpm_cnn_set_params", [&] {
        Net* net = (Net*)h;
        if (!net || !in) throw fpm::HipError{hipErrorInvalidValue, "null argument"};
        FPM_HIP(hipMemcpyAsync(net->params.p, in, net->nparam * 4, hipMemcpyHostToDevice, net->st));
        to_bf16(net->params.as<float>(), net->wbf.as<bf16_t>(), net->nparam, net->st);
        fill_f32(net->m.as<float>(), net->nparam, 0.0f, net->st);
        fill_f32(net->v.as<float>(), net->nparam, 0.0f, net->st);
        net->step = 0;
        FPM_HIP(hipStreamSynchronize(net->st));
    });
}

// The network's stream is its own (non-blocking); its inputs are often
// written by work queued on the null stream just before the call (torch's
// default stream): wait for that work first.
static void after_null_stream(hipStream_t st) {
    hipEvent_t e = nullptr;
    FPM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipError_t r = hipEventRecord(e, nullptr);
    if (r == hipSuccess) r = hipStreamWaitEvent(st, e, 0);
    (void)hipEventDestroy(e);
    FPM_HIP(r);
}

int fpm_cnn_forward(fpm_cnn* h, const float* feats, int64_t batch, float* out, float* pupil_out) {
    return guarded("fpm_cnn_forward", [&] {
        Net* net = (Net*)h;
        if (!net || !feats || !out) throw fpm::HipError{hipErrorInvalidValue, "null argument"};
        if (batch < 1 || batch > net->d.max_batch) throw fpm::HipError{hipErrorInvalidValue, "batch outside [1, max_batch]"};
        const int B = (int)batch;
        if (pupil_out && net->K == 0)
            throw fpm::HipError{hipErrorInvalidValue, "pupil_out given to a network without a pupil head"};
        after_null_stream(net->st);
        load_features(*net, feats, B);
        forward(*net, B, pupil_out);
        head_shuffle(net->head_f.as<float>(), out, B, net->d.n, net->d.n, net->d.up, net->st);
        FPM_HIP(hipStreamSynchronize(net->st));
    });
}

int fpm_cnn_gradient(fpm_cnn* h, const float* feats, const float* targets, const float* pupil_targets, int64_t batch,
                     double* loss_out, float* grad_out) {
    return guarded("fpm_cnn_gradient", [&] {
        Net* net = (Net*)h;
        check_call(net, feats, targets, pupil_targets, batch);
        after_null_stream(net->st);
        const double loss = loss_and_grad(*net, feats, targets, pupil_targets, (int)batch);
        if (loss_out) *loss_out = loss;
        if (grad_out) {
            FPM_HIP(hipMemcpyAsync(grad_out, net->grads.p, net->nparam * 4, hipMemcpyDeviceToHost, net->st));
            FPM_HIP(hipStreamSynchronize(net->st));
        }
    });
}

int fpm_cnn_train_step(fpm_cnn* h, const float* feats, const float* targets, const float* pupil_targets,
                       int64_t batch, float lr, double* loss_out) {
    return guarded("fpm_cnn_train_step", [&] {
        Net* net = (Net*)h;
        check_call(net, feats, targets, pupil_targets, batch);
        after_null_stream(net->st);
        const double loss = loss_and_grad(*net, feats, targets, pupil_targets, (int)batch);
        net->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        AdamArgs a{lr, b1, b2, 1e-8f, (float)(1.0 / (1.0 - std::pow((double)b1, (double)net->step))),
                   (float)(1.0 / (1.0 - std::pow((double)b2, (double)net->step)))};
        adam_update(net->params.as<float>(), net->m.as<float>(), net->v.as<float>(), net->grads.as<float>(),
                    net->nparam, a, net->st);
        to_bf16(net->params.as<float>(), net->wbf.as<bf16_t>(), net->nparam, net->st);
        FPM_HIP(hipStreamSynchronize(net->st));
        if (loss_out) *loss_out = loss;
    });
}

int fpm_parallax_shifts(const float* planes, int64_t batch, int64_t planes_per_sample, int64_t rows, int64_t cols,
                        int32_t ref, const int32_t* which, int32_t count, int32_t iters, float* shifts, void* stream) {
    return guarded("fpm_parallax_shifts", [&] {
        if (!planes || !which || !shifts) throw fpm::HipError{hipErrorInvalidValue, "null argument"};
        if (batch < 1 || batch > 65535 || planes_per_sample < 1 || planes_per_sample > (1 << 20) || rows < 3 ||
            cols < 3 || rows * cols >= ((int64_t)1 << 31))
            throw fpm::HipError{hipErrorInvalidValue, "batch in [1, 65535], planes of at least 3 x 3"};
        hipStream_t st = (hipStream_t)stream;
        parallax_shifts(planes, (int)batch, (int)planes_per_sample, (int)rows, (int)cols, ref, which, count, iters,
                        shifts, st);
        FPM_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
