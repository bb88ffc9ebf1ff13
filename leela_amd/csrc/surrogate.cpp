// surrogate.cpp — the learned playout surrogate behind the C ABI: FP32 master
// parameters with Adam moments, BF16 packed weights, the forward pass (kept
// activations for training), the backward pass, and checkpoints.
//
// HBM layout per surrogate (max_batch B):
//   params / grads / Adam m / Adam v   4 x 484,930 FP32
//   packed weights                     13 forward + 12 backward-data convs
//   activations                        13 x B x 361 x 64 BF16 (all kept)
//   conv-gradient maps                 3 x B x 361 x 64 BF16 (rotating)
//   head saves s/f/t                   B x (361 + 64 + 1) FP32
//   partial sums                       max of the wgrad and head partials
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "leela_amd.h"
#include "lz_net.h"
#include "lz_surrogate.h"

namespace {

using namespace lz::sg;

constexpr uint32_t kMagic = 0x47535a4c;  // "LZSG"
constexpr uint32_t kVersion = 1;

struct Layout {
    size_t w[kNConv], b[kNConv], head, total;
    Layout() {
        size_t off = 0;
        for (int l = 0; l < kNConv; l++) {
            w[l] = off;
            off += conv_weights(l);
            b[l] = off;
            off += kCh;
        }
        head = off;
        total = off + kHeadParams;
    }
};

const Layout &layout() {
    static const Layout L;
    return L;
}

}  // namespace

struct lz_surrogate {
    int device = 0;
    int max_batch = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string err;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wf[kNConv] = {}, *wb[kNConv] = {};
    uint16_t *act = nullptr, *e[3] = {};
    uint16_t *ekeep = nullptr;   // all 13 gradient maps (inspection mode)
    float *s = nullptr, *f = nullptr, *t = nullptr, *y = nullptr;
    float *loss = nullptr, *part = nullptr;
    uint8_t *planes = nullptr;   // staging of the host entry points
    float *win = nullptr;
    hipEvent_t ev = nullptr;
    long long step = 0;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    bool packed = false;

    int fail(int code, const std::string &what) {
        err = what;
        return code;
    }
    int hip(hipError_t e, const char *what) {
        if (e == hipSuccess) return LZ_OK;
        return fail(LZ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
    size_t map_elems() const { return (size_t)max_batch * kP * kCh; }
    uint16_t *act_of(int l) const { return act + (size_t)l * map_elems(); }
    // Gradient map at the pre-activation of conv l. Three rotating maps
    // suffice: computing map l-1 reads maps l and l+1.
    uint16_t *emap(int l) const {
        return ekeep ? ekeep + (size_t)l * map_elems() : e[l % 3];
    }
};

namespace {

template <typename T>
hipError_t dalloc(T **ptr, size_t count) {
    return hipMalloc((void **)ptr, count * sizeof(T) + 256);
}

void release(lz_surrogate *s) {
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    void *bufs[] = {s->p, s->g, s->m, s->v, s->act, s->e[0], s->e[1], s->e[2],
                    s->ekeep, s->s, s->f, s->t, s->y, s->loss, s->part,
                    s->planes, s->win};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < kNConv; l++) {
        if (s->wf[l]) (void)hipFree(s->wf[l]);
        if (s->wb[l]) (void)hipFree(s->wb[l]);
    }
    if (s->ev) (void)hipEventDestroy(s->ev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

int repack(lz_surrogate *s, hipStream_t st) {
    const Layout &L = layout();
    for (int l = 0; l < kNConv; l++) {
        int rc = s->hip(pack_conv(s->p + L.w[l], s->wf[l], conv_cin(l), false,
                                  st), "pack");
        if (rc) return rc;
        if (l > 0) {
            rc = s->hip(pack_conv(s->p + L.w[l], s->wb[l], kCh, true, st),
                        "pack");
            if (rc) return rc;
        }
    }
    s->packed = true;
    return LZ_OK;
}

// Forward over n boards of device planes; keeps every activation.
int forward(lz_surrogate *s, int n, const uint8_t *planes, float *win,
            hipStream_t st) {
    const Layout &L = layout();
    if (!s->packed) {
        int rc = repack(s, st);
        if (rc) return rc;
    }
    for (int l = 0; l < kNConv; l++) {
        const void *in = l == 0 ? (const void *)planes : (const void *)s->act_of(l - 1);
        const uint16_t *res = (l >= 2 && l % 2 == 0) ? s->act_of(l - 2) : nullptr;
        int rc = s->hip(conv(in, conv_cin(l), s->wf[l], s->p + L.b[l], res,
                             nullptr, s->act_of(l), n, false, st), "conv");
        if (rc) return rc;
    }
    return s->hip(head_fwd(s->act_of(kNConv - 1), s->p + L.head, s->s, s->f,
                           s->t, win, n, st), "head_fwd");
}

// Backward of the last forward into s->g (written, not accumulated); the
// mean loss lands in s->loss (device). Map l is the gradient at the
// pre-activation of conv l: the second conv of a block (even l >= 2) passes
// map l to its first conv, and the first conv (odd l) adds map l + 1 (the
// skip) before masking by the block input.
int backward(lz_surrogate *s, int n, const uint8_t *planes, const float *y,
             hipStream_t st) {
    const Layout &L = layout();
    const int last = kNConv - 1;
    int rc = s->hip(head_bwd(s->act_of(last), s->p + L.head, s->s, s->f, s->t,
                             y, s->emap(last), s->part, s->g + L.head, s->loss,
                             n, st), "head_bwd");
    for (int l = last; !rc && l >= 1; l--) {
        rc = s->hip(wgrad(s->act_of(l - 1), kCh, s->emap(l), s->part,
                          s->g + L.w[l], s->g + L.b[l], n, st), "wgrad");
        if (rc) return rc;
        const uint16_t *res = (l % 2 == 1) ? s->emap(l + 1) : nullptr;
        rc = s->hip(conv(s->emap(l), kCh, s->wb[l], nullptr, res,
                         s->act_of(l - 1), s->emap(l - 1), n, true, st),
                    "conv_bwd");
    }
    if (rc) return rc;
    return s->hip(wgrad(planes, 32, s->emap(0), s->part, s->g + L.w[0],
                        s->g + L.b[0], n, st), "wgrad");
}

// Orders `to` after the work queued so far on `from`.
int join(lz_surrogate *s, hipStream_t from, hipStream_t to) {
    if (from == to) return LZ_OK;
    int rc = s->hip(hipEventRecord(s->ev, from), "event record");
    if (!rc) rc = s->hip(hipStreamWaitEvent(to, s->ev, 0), "stream wait");
    return rc;
}

int check(lz_surrogate *s, int n) {
    if (!s) return LZ_ERR_ARG;
    if (n <= 0 || n > s->max_batch)
        return s->fail(LZ_ERR_ARG, "batch size outside 1..max_batch");
    if (hipSetDevice(s->device) != hipSuccess)
        return s->fail(LZ_ERR_HIP, "hipSetDevice failed");
    return LZ_OK;
}

}  // namespace

extern "C" {

int lz_surrogate_create(lz_surrogate **out, int device, int max_batch,
                        uint32_t seed) {
    if (!out || max_batch <= 0) return LZ_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return LZ_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return LZ_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return LZ_ERR_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return LZ_ERR_UNSUPPORTED;
    lz_surrogate *s = new (std::nothrow) lz_surrogate();
    if (!s) return LZ_ERR_NOMEM;
    s->device = device;
    s->max_batch = max_batch;
    const Layout &L = layout();
    const size_t B = (size_t)max_batch;
    size_t part = head_part_floats(max_batch);
    for (int cin : {32, kCh}) {
        const size_t w = wgrad_part_floats(cin, max_batch);
        if (w > part) part = w;
    }
    bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&s->ev, hipEventDisableTiming) == hipSuccess;
    ok = ok && dalloc(&s->p, L.total) == hipSuccess &&
         dalloc(&s->g, L.total) == hipSuccess &&
         dalloc(&s->m, L.total) == hipSuccess &&
         dalloc(&s->v, L.total) == hipSuccess &&
         dalloc(&s->act, kNConv * s->map_elems()) == hipSuccess &&
         dalloc(&s->s, B * kP) == hipSuccess &&
         dalloc(&s->f, B * kHid) == hipSuccess &&
         dalloc(&s->t, B) == hipSuccess && dalloc(&s->y, B) == hipSuccess &&
         dalloc(&s->win, B) == hipSuccess && dalloc(&s->loss, 1) == hipSuccess &&
         dalloc(&s->part, part) == hipSuccess &&
         dalloc(&s->planes, B * 32 * kP) == hipSuccess;
    for (int i = 0; ok && i < 3; i++) ok = dalloc(&s->e[i], s->map_elems()) == hipSuccess;
    for (int l = 0; ok && l < kNConv; l++) {
        ok = dalloc(&s->wf[l], (size_t)conv_weights(l)) == hipSuccess;
        if (ok && l > 0) ok = dalloc(&s->wb[l], (size_t)conv_weights(l)) == hipSuccess;
    }
    if (!ok) {
        release(s);
        delete s;
        return LZ_ERR_NOMEM;
    }
    // He-normal initialisation from the seed, biases zero. The last conv of
    // every residual block and the output layer F2 start at zero, so each
    // block starts as the identity and the output at t = 0 (win rate 0.5):
    // without normalisation layers the unscaled blocks compound and Adam at
    // lr 1e-3 drove the tanh into saturation within the first steps.
    std::vector<float> h(L.total, 0.f);
    for (int l = 0; l < kNConv; l++) {
        if (l >= 2 && l % 2 == 0) continue;
        lz::normal_fill(seed * 100u + (uint32_t)l, conv_weights(l),
                        std::sqrt(2.f / (conv_cin(l) * 9)), h.data() + L.w[l]);
    }
    float *hd = h.data() + L.head;
    lz::normal_fill(seed * 100u + 50u, kCh, std::sqrt(2.f / kCh), hd);
    lz::normal_fill(seed * 100u + 51u, (size_t)kHid * kP, std::sqrt(2.f / kP),
                    hd + kCh + 1);
    ok = hipMemcpy(s->p, h.data(), L.total * sizeof(float),
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(s->m, 0, L.total * sizeof(float)) == hipSuccess &&
         hipMemset(s->v, 0, L.total * sizeof(float)) == hipSuccess;
    if (!ok) {
        release(s);
        delete s;
        return LZ_ERR_HIP;
    }
    *out = s;
    return LZ_OK;
}

void lz_surrogate_destroy(lz_surrogate *s) {
    if (!s) return;
    release(s);
    delete s;
}

int lz_surrogate_num_params(void) { return (int)layout().total; }

const char *lz_surrogate_last_error(const lz_surrogate *s) {
    return s ? s->err.c_str() : "null surrogate";
}

int lz_surrogate_get_params(lz_surrogate *s, float *params) {
    if (!s || !params) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    (void)hipSetDevice(s->device);
    int rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    if (rc) return rc;
    return s->hip(hipMemcpy(params, s->p, layout().total * sizeof(float),
                            hipMemcpyDeviceToHost), "copy params");
}

int lz_surrogate_set_params(lz_surrogate *s, const float *params) {
    if (!s || !params) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    (void)hipSetDevice(s->device);
    const size_t bytes = layout().total * sizeof(float);
    int rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    if (!rc) rc = s->hip(hipMemcpy(s->p, params, bytes, hipMemcpyHostToDevice), "copy");
    if (!rc) rc = s->hip(hipMemset(s->m, 0, bytes), "memset");
    if (!rc) rc = s->hip(hipMemset(s->v, 0, bytes), "memset");
    s->step = 0;
    s->packed = false;
    return rc;
}

int lz_surrogate_set_adam(lz_surrogate *s, float beta1, float beta2, float eps) {
    if (!s || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) ||
        !(eps > 0.f))
        return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    s->beta1 = beta1;
    s->beta2 = beta2;
    s->eps = eps;
    return LZ_OK;
}

long long lz_surrogate_step(const lz_surrogate *s) { return s ? s->step : -1; }

int lz_surrogate_forward_device(lz_surrogate *s, int n, const uint8_t *planes,
                                float *winrate, void *stream) {
    int rc = check(s, n);
    if (rc) return rc;
    if (!planes || !winrate) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    rc = join(s, s->stream, st);
    if (!rc) rc = forward(s, n, planes, winrate, st);
    if (!rc) rc = join(s, st, s->stream);
    return rc;
}

int lz_surrogate_predict(lz_surrogate *s, int n, const uint8_t *planes,
                         float *winrate) {
    int rc = check(s, n);
    if (rc) return rc;
    if (!planes || !winrate) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    rc = s->hip(hipMemcpyAsync(s->planes, planes, (size_t)n * 32 * kP,
                               hipMemcpyHostToDevice, s->stream), "copy planes");
    if (!rc) rc = forward(s, n, s->planes, s->win, s->stream);
    if (!rc) rc = s->hip(hipMemcpyAsync(winrate, s->win, n * sizeof(float),
                                        hipMemcpyDeviceToHost, s->stream), "copy out");
    if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    return rc;
}

int lz_surrogate_gradients(lz_surrogate *s, int n, const uint8_t *planes,
                           const float *targets, float *grad, float *loss) {
    int rc = check(s, n);
    if (rc) return rc;
    if (!planes || !targets) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    rc = s->hip(hipMemcpyAsync(s->planes, planes, (size_t)n * 32 * kP,
                               hipMemcpyHostToDevice, s->stream), "copy planes");
    if (!rc) rc = s->hip(hipMemcpyAsync(s->y, targets, n * sizeof(float),
                                        hipMemcpyHostToDevice, s->stream), "copy y");
    if (!rc) rc = forward(s, n, s->planes, s->win, s->stream);
    if (!rc) rc = backward(s, n, s->planes, s->y, s->stream);
    if (!rc && grad)
        rc = s->hip(hipMemcpyAsync(grad, s->g, layout().total * sizeof(float),
                                   hipMemcpyDeviceToHost, s->stream), "copy grad");
    if (!rc && loss)
        rc = s->hip(hipMemcpyAsync(loss, s->loss, sizeof(float),
                                   hipMemcpyDeviceToHost, s->stream), "copy loss");
    if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    return rc;
}

int lz_surrogate_train_step_device(lz_surrogate *s, int n,
                                   const uint8_t *planes, const float *targets,
                                   float lr, float *loss, void *stream) {
    int rc = check(s, n);
    if (rc) return rc;
    if (!planes || !targets || !(lr >= 0.f)) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
    rc = join(s, s->stream, st);
    if (!rc) rc = forward(s, n, planes, nullptr, st);
    if (!rc) rc = backward(s, n, planes, targets, st);
    if (!rc) {
        s->step++;
        const double c1 = 1.0 - std::pow((double)s->beta1, (double)s->step);
        const double c2 = 1.0 - std::pow((double)s->beta2, (double)s->step);
        rc = s->hip(adam(s->p, s->g, s->m, s->v, (int)layout().total, lr,
                         s->beta1, s->beta2, s->eps, (float)c1, (float)c2, st),
                    "adam");
    }
    if (!rc) rc = repack(s, st);
    if (!rc && loss)
        rc = s->hip(hipMemcpyAsync(loss, s->loss, sizeof(float),
                                   hipMemcpyDeviceToHost, st), "copy loss");
    if (!rc) rc = join(s, st, s->stream);
    if (!rc && loss) rc = s->hip(hipStreamSynchronize(st), "sync");
    return rc;
}

int lz_surrogate_keep_maps(lz_surrogate *s, int on) {
    if (!s) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    (void)hipSetDevice(s->device);
    int rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    if (rc) return rc;
    if (on && !s->ekeep) {
        if (dalloc(&s->ekeep, kNConv * s->map_elems()) != hipSuccess) {
            s->ekeep = nullptr;
            return s->fail(LZ_ERR_NOMEM, "gradient maps");
        }
    } else if (!on && s->ekeep) {
        rc = s->hip(hipFree(s->ekeep), "free");
        s->ekeep = nullptr;
    }
    return rc;
}

int lz_surrogate_read_map(lz_surrogate *s, int kind, int layer, int n,
                          uint16_t *out) {
    int rc = check(s, n);
    if (rc) return rc;
    if (!out || layer < 0 || layer >= kNConv || kind < 0 || kind > 1)
        return LZ_ERR_ARG;
    if (kind == 1 && !s->ekeep)
        return s->fail(LZ_ERR_STATE, "gradient maps are not kept");
    std::lock_guard<std::mutex> g(s->mu);
    const uint16_t *src = kind == 0 ? s->act_of(layer) : s->emap(layer);
    rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    if (!rc)
        rc = s->hip(hipMemcpy(out, src, (size_t)n * kP * kCh * sizeof(uint16_t),
                              hipMemcpyDeviceToHost), "copy map");
    return rc;
}

int lz_surrogate_read_head(lz_surrogate *s, int n, float *sv, float *fv,
                           float *tv) {
    int rc = check(s, n);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    if (!rc && sv)
        rc = s->hip(hipMemcpy(sv, s->s, (size_t)n * kP * sizeof(float),
                              hipMemcpyDeviceToHost), "copy s");
    if (!rc && fv)
        rc = s->hip(hipMemcpy(fv, s->f, (size_t)n * kHid * sizeof(float),
                              hipMemcpyDeviceToHost), "copy f");
    if (!rc && tv)
        rc = s->hip(hipMemcpy(tv, s->t, (size_t)n * sizeof(float),
                              hipMemcpyDeviceToHost), "copy t");
    return rc;
}

int lz_surrogate_save(lz_surrogate *s, const char *path) {
    if (!s || !path) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    (void)hipSetDevice(s->device);
    const size_t N = layout().total;
    std::vector<float> buf(3 * N);
    int rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    float *dev[3] = {s->p, s->m, s->v};
    for (int i = 0; !rc && i < 3; i++)
        rc = s->hip(hipMemcpy(buf.data() + i * N, dev[i], N * sizeof(float),
                              hipMemcpyDeviceToHost), "copy");
    if (rc) return rc;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return s->fail(LZ_ERR_ARG, std::string("cannot write ") + path);
    const uint32_t hdr[3] = {kMagic, kVersion, (uint32_t)N};
    const int64_t step = s->step;
    const float hyper[3] = {s->beta1, s->beta2, s->eps};
    bool ok = std::fwrite(hdr, sizeof(hdr), 1, fp) == 1 &&
              std::fwrite(&step, sizeof(step), 1, fp) == 1 &&
              std::fwrite(hyper, sizeof(hyper), 1, fp) == 1 &&
              std::fwrite(buf.data(), sizeof(float), buf.size(), fp) == buf.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? LZ_OK : s->fail(LZ_ERR_ARG, std::string("short write to ") + path);
}

int lz_surrogate_load(lz_surrogate *s, const char *path) {
    if (!s || !path) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    (void)hipSetDevice(s->device);
    const size_t N = layout().total;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return s->fail(LZ_ERR_ARG, std::string("cannot read ") + path);
    uint32_t hdr[3];
    int64_t step = 0;
    float hyper[3];
    std::vector<float> buf(3 * N);
    bool ok = std::fread(hdr, sizeof(hdr), 1, fp) == 1 && hdr[0] == kMagic &&
              hdr[1] == kVersion && hdr[2] == N &&
              std::fread(&step, sizeof(step), 1, fp) == 1 &&
              std::fread(hyper, sizeof(hyper), 1, fp) == 1 &&
              std::fread(buf.data(), sizeof(float), buf.size(), fp) == buf.size();
    std::fclose(fp);
    if (!ok) return s->fail(LZ_ERR_ARG, std::string("not a surrogate checkpoint: ") + path);
    int rc = s->hip(hipStreamSynchronize(s->stream), "sync");
    float *dev[3] = {s->p, s->m, s->v};
    for (int i = 0; !rc && i < 3; i++)
        rc = s->hip(hipMemcpy(dev[i], buf.data() + i * N, N * sizeof(float),
                              hipMemcpyHostToDevice), "copy");
    if (rc) return rc;
    s->step = step;
    s->beta1 = hyper[0];
    s->beta2 = hyper[1];
    s->eps = hyper[2];
    s->packed = false;
    return LZ_OK;
}

}  // extern "C"
