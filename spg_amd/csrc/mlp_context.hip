// Host side of the learned warm start network (include/spg_vbd_learned.h): parameters, training
// loop (forward, loss, backward, Adam, all on the device through mlp_kernels.hip) and inference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.h"
#include "mlp_device.h"
#include "spg_vbd_learned.h"

using namespace spgvbd;

namespace {

thread_local std::string g_mlpLastError;

constexpr int kLayerIn[kMlpLayers] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden};
constexpr int kLayerOut[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpOutPad};
constexpr int kFlatIn[kMlpLayers] = {kMlpFeatures, kMlpHidden, kMlpHidden, kMlpHidden};
constexpr int kFlatOut[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};

uint64_t splitmix64(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return static_cast<double>(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int64_t roundUp(int64_t n, int64_t m) { return (n + m - 1) / m * m; }

}  // namespace

struct spg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    int64_t wOff[kMlpLayers]{}, bOff[kMlpLayers]{}, total = 0;  // padded master layout
    DevBuf<float> P, G, M, V;                                   // parameters, gradients, Adam moments
    DevBuf<__bf16> wb[kMlpLayers], wtb[kMlpLayers];              // BF16 W (out x in) and W^T (in x out)
    DevBuf<float> mean, invStd;
    DevBuf<unsigned char> frag;  // weights in the fused inference kernel's operand order
    int maxBlocks = 512;         // persistent workgroups of the fused inference kernel
    std::vector<float> hMean, hStd;
    int64_t adamT = 0;
    int32_t historyBase = 0;  // 1: the output is the change of the previous substep's offset (see spg_mlp_set_history_base)

    // workspaces
    int64_t trainRows = 0;
    DevBuf<__bf16> X, XT, H[2], HT[3], dY, dYT, dZb[2], dZT;
    DevBuf<float> T, Z[3], Yf, dYf, dZf;
    DevBuf<double> loss;
    DevBuf<int32_t> index;
    DevBuf<float> dataF, dataT;

    ~spg_mlp()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void init(uint64_t seed)
    {
        total = 0;
        for (int l = 0; l < kMlpLayers; ++l) {
            wOff[l] = total;
            total += static_cast<int64_t>(kLayerOut[l]) * kLayerIn[l];
            bOff[l] = total;
            total += kLayerOut[l];
        }
        P.alloc(total);
        G.alloc(total);
        M.alloc(total);
        V.alloc(total);
        for (int l = 0; l < kMlpLayers; ++l) {
            wb[l].alloc(static_cast<size_t>(kLayerOut[l]) * kLayerIn[l]);
            wtb[l].alloc(static_cast<size_t>(kLayerOut[l]) * kLayerIn[l]);
        }
        mean.alloc(kMlpFeatures);
        invStd.alloc(kMlpFeatures);
        frag.alloc(mlpFragBytes());
        int cus = 0;
        hipCheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
        maxBlocks = 2 * std::max(cus, 1);  // two 80 KB-LDS workgroups fit a CU
        std::vector<float> flat(SPG_MLP_NUM_PARAMS, 0.f);
        uint64_t s = seed;
        int64_t k = 0;
        for (int l = 0; l < kMlpLayers; ++l) {
            const double a = std::sqrt(6.0 / (kFlatIn[l] + kFlatOut[l]));
            for (int i = 0; i < kFlatOut[l] * kFlatIn[l]; ++i) flat[k++] = static_cast<float>((2.0 * uniform01(s) - 1.0) * a);
            k += kFlatOut[l];  // zero biases
        }
        setParams(flat.data());
        std::vector<float> m(kMlpFeatures, 0.f), sd(kMlpFeatures, 1.f);
        setNormalization(m.data(), sd.data());
    }

    // flat (unpadded) <-> padded master layout
    template <class F>
    void forEachParam(F &&f) const
    {
        int64_t k = 0;
        for (int l = 0; l < kMlpLayers; ++l) {
            for (int r = 0; r < kFlatOut[l]; ++r)
                for (int c = 0; c < kFlatIn[l]; ++c) f(k++, wOff[l] + static_cast<int64_t>(r) * kLayerIn[l] + c);
            for (int r = 0; r < kFlatOut[l]; ++r) f(k++, bOff[l] + r);
        }
    }

    void setParams(const float *flat)
    {
        std::vector<float> padded(total, 0.f);
        forEachParam([&](int64_t k, int64_t p) { padded[p] = flat[k]; });
        for (float v : padded)
            if (!std::isfinite(v)) throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite parameter");
        hipCheck(hipMemcpyAsync(P.p, padded.data(), sizeof(float) * total, hipMemcpyHostToDevice, stream), "upload");
        hipCheck(hipMemsetAsync(M.p, 0, sizeof(float) * total, stream), "memset");
        hipCheck(hipMemsetAsync(V.p, 0, sizeof(float) * total, stream), "memset");
        adamT = 0;
        refreshWeights();
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    std::vector<float> downloadFlat(const float *dev)
    {
        std::vector<float> padded(total), flat(SPG_MLP_NUM_PARAMS);
        hipCheck(hipMemcpyAsync(padded.data(), dev, sizeof(float) * total, hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        forEachParam([&](int64_t k, int64_t p) { flat[k] = padded[p]; });
        return flat;
    }

    void setNormalization(const float *m, const float *sd)
    {
        std::vector<float> inv(kMlpFeatures);
        for (int i = 0; i < kMlpFeatures; ++i) {
            if (!(sd[i] > 0.f) || !std::isfinite(sd[i]) || !std::isfinite(m[i]))
                throw VbdError(SPG_VBD_ERR_INVALID, "normalisation needs finite means and positive standard deviations");
            inv[i] = 1.f / sd[i];
        }
        hMean.assign(m, m + kMlpFeatures);
        hStd.assign(sd, sd + kMlpFeatures);
        hipCheck(hipMemcpyAsync(mean.p, m, sizeof(float) * kMlpFeatures, hipMemcpyHostToDevice, stream), "upload");
        hipCheck(hipMemcpyAsync(invStd.p, inv.data(), sizeof(float) * kMlpFeatures, hipMemcpyHostToDevice, stream),
                 "upload");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void refreshWeights()
    {
        for (int l = 0; l < kMlpLayers; ++l)
            hipCheck(launchWeightCopies(P.p + wOff[l], kLayerOut[l], kLayerIn[l], wb[l].p, wtb[l].p, stream),
                     "weight copies");
        hipCheck(launchMlpFrag(P.p + wOff[0], P.p + wOff[1], P.p + wOff[2], P.p + wOff[3], frag.p, stream),
                 "fragment launch");
    }

    void reserveTrain(int64_t rows)
    {
        if (rows <= trainRows) return;
        X.alloc(rows * kMlpIn);
        XT.alloc(rows * kMlpIn);
        T.alloc(rows * kMlpOut);
        for (int i = 0; i < 2; ++i) H[i].alloc(rows * kMlpHidden);
        for (int i = 0; i < 3; ++i) {
            HT[i].alloc(rows * kMlpHidden);
            Z[i].alloc(rows * kMlpHidden);
        }
        Yf.alloc(rows * kMlpInferCols);
        dYf.alloc(rows * kMlpOutPad);
        dY.alloc(rows * kMlpOutPad);
        dYT.alloc(rows * kMlpOutPad);
        for (int i = 0; i < 2; ++i) dZb[i].alloc(rows * kMlpHidden);
        dZT.alloc(rows * kMlpHidden);
        dZf.alloc(rows * kMlpHidden);
        loss.alloc(1);
        trainRows = rows;
    }

    void gemm(const __bf16 *A, int lda, const __bf16 *B, int ldb, int64_t Mr, int N, int K, const GemmEpilogue &ep)
    {
        hipCheck(launchGemmNT(A, lda, B, ldb, static_cast<int>(Mr), N, K, ep, stream), "gemm launch");
    }

    // Forward pass over `rows` packed rows in X; y (FP32, stride ldy >= 16) receives the outputs. With
    // `train`, keeps Z (pre-activations) and transposed activations for the backward pass.
    void forward(int64_t rows, bool train, float *y, int ldy)
    {
        const __bf16 *A = X.p;
        int lda = kMlpIn;
        for (int l = 0; l < 3; ++l) {
            GemmEpilogue ep;
            ep.bias = P.p + bOff[l];
            ep.act = 1;
            ep.hOut = H[l & 1].p;
            ep.ldh = kMlpHidden;
            if (train) {
                ep.zOut = Z[l].p;
                ep.ldz = kMlpHidden;
                ep.hTOut = HT[l].p;
                ep.ldht = static_cast<int>(rows);
            }
            gemm(A, lda, wb[l].p, kLayerIn[l], rows, kMlpHidden, kLayerIn[l], ep);
            A = H[l & 1].p;
            lda = kMlpHidden;
        }
        GemmEpilogue ep;
        ep.bias = P.p + bOff[3];
        ep.zOut = y;
        ep.ldz = ldy;
        gemm(A, lda, wb[3].p, kMlpHidden, rows, kMlpInferCols, kMlpHidden, ep);
    }

    // Gradient of the batch loss into G, given the forward pass of `rows` rows (n real) and targets T.
    void backward(int64_t rows, int n)
    {
        const int r = static_cast<int>(rows);
        hipCheck(launchLossGrad(Yf.p, kMlpInferCols, T.p, kMlpOut, n, r, kMlpOutPad, dYf.p, dY.p, dYT.p, loss.p, stream),
                 "loss launch");
        GemmEpilogue e;
        // layer 4: dW4 = dY^T H3, db4 = colsum dY
        e.zOut = G.p + wOff[3];
        e.ldz = kMlpHidden;
        gemm(dYT.p, r, HT[2].p, r, kMlpOutPad, kMlpHidden, r, e);
        hipCheck(launchColSum(dYf.p, kMlpOutPad, r, kMlpOutPad, G.p + bOff[3], stream), "colsum");
        // dZ3 = (dY W4) . silu'(Z3)
        const __bf16 *dIn = dY.p;
        int ldIn = kMlpOutPad, kIn = kMlpOutPad;
        for (int l = 2; l >= 0; --l) {
            GemmEpilogue d;
            d.dactZ = Z[l].p;
            d.lddz = kMlpHidden;
            d.zOut = dZf.p;
            d.ldz = kMlpHidden;
            if (l > 0) {
                d.hOut = dZb[l & 1].p;
                d.ldh = kMlpHidden;
            }
            d.hTOut = dZT.p;
            d.ldht = r;
            gemm(dIn, ldIn, wtb[l + 1].p, kLayerOut[l + 1], rows, kMlpHidden, kIn, d);
            // dW_l = dZ_l^T H_{l-1} (X for the first layer), db_l = colsum dZ_l
            GemmEpilogue w;
            w.zOut = G.p + wOff[l];
            w.ldz = kLayerIn[l];
            gemm(dZT.p, r, l > 0 ? HT[l - 1].p : XT.p, r, kMlpHidden, kLayerIn[l], r, w);
            hipCheck(launchColSum(dZf.p, kMlpHidden, r, kMlpHidden, G.p + bOff[l], stream), "colsum");
            dIn = dZb[l & 1].p;
            ldIn = kMlpHidden;
            kIn = kMlpHidden;
        }
    }

    // Fused inference on device features (n x 24) into y (stride ldy >= 3), ordered on the caller's
    // stream `s` (runs on the network's stream between two events).
    void infer(const float *features, int64_t n, float *y, int ldy, hipStream_t s)
    {
        if (n > (int64_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many rows");
        hipEvent_t ev;
        hipCheck(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        hipCheck(hipEventRecord(ev, s), "hipEventRecord");
        hipCheck(hipStreamWaitEvent(stream, ev, 0), "hipStreamWaitEvent");
        hipCheck(launchMlpInfer(features, static_cast<int>(n), mean.p, invStd.p, frag.p, P.p + bOff[0], P.p + bOff[1],
                                P.p + bOff[2], P.p + bOff[3], y, ldy, maxBlocks, stream),
                 "inference launch");
        hipCheck(hipEventRecord(ev, stream), "hipEventRecord");
        hipCheck(hipStreamWaitEvent(s, ev, 0), "hipStreamWaitEvent");
        hipCheck(hipEventDestroy(ev), "hipEventDestroy");
    }
};

namespace spgvbd {

void mlpInferDevice(spg_mlp *net, const float *features, int64_t n, float *y, int32_t ldy, hipStream_t stream)
{
    if (ldy < kMlpOut) throw VbdError(SPG_VBD_ERR_INVALID, "output stride below 3");
    hipCheck(hipSetDevice(net->device), "hipSetDevice");
    net->infer(features, n, y, ldy, stream);
}

void mlpInferAfter(spg_mlp *net, const float *features, int64_t n, float *y, int32_t ldy, hipEvent_t ready,
                   hipEvent_t done)
{
    if (ldy < kMlpOut) throw VbdError(SPG_VBD_ERR_INVALID, "output stride below 3");
    if (n > (int64_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many rows");
    hipCheck(hipSetDevice(net->device), "hipSetDevice");
    hipCheck(hipStreamWaitEvent(net->stream, ready, 0), "hipStreamWaitEvent");
    hipCheck(launchMlpInfer(features, static_cast<int>(n), net->mean.p, net->invStd.p, net->frag.p, net->P.p + net->bOff[0],
                            net->P.p + net->bOff[1], net->P.p + net->bOff[2], net->P.p + net->bOff[3], y, ldy,
                            net->maxBlocks, net->stream),
             "inference launch");
    hipCheck(hipEventRecord(done, net->stream), "hipEventRecord");
}

int mlpDevice(const spg_mlp *net) { return net->device; }
int mlpHistoryBase(const spg_mlp *net) { return net->historyBase; }

}  // namespace spgvbd

namespace {

template <class F>
int guardedMlp(spg_mlp *net, F &&f)
{
    try {
        if (net) hipCheck(hipSetDevice(net->device), "hipSetDevice");
        f();
        if (net) net->err.clear();
        return SPG_VBD_OK;
    } catch (const VbdError &e) {
        if (net) net->err = e.what();
        g_mlpLastError = e.what();
        return e.code;
    } catch (const std::exception &e) {
        if (net) net->err = e.what();
        g_mlpLastError = e.what();
        return SPG_VBD_ERR_INVALID;
    }
}

void requireFinite(const float *p, int64_t n, const char *what)
{
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) throw VbdError(SPG_VBD_ERR_NONFINITE, std::string("non-finite value in ") + what);
}

}  // namespace

extern "C" {

int spg_mlp_create(spg_mlp **out, int32_t device, uint64_t seed)
{
    if (!out) {
        g_mlpLastError = "null argument";
        return SPG_VBD_ERR_INVALID;
    }
    *out = nullptr;
    spg_mlp *net = new spg_mlp();
    net->device = device;
    const int rc = guardedMlp(nullptr, [&] {
        int count = 0;
        hipCheck(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) throw VbdError(SPG_VBD_ERR_INVALID, "invalid device");
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipCheck(hipStreamCreate(&net->stream), "hipStreamCreate");
        net->init(seed);
    });
    if (rc != SPG_VBD_OK) {
        delete net;
        return rc;
    }
    *out = net;
    return SPG_VBD_OK;
}

void spg_mlp_destroy(spg_mlp *net)
{
    if (!net) return;
    (void)hipSetDevice(net->device);
    delete net;
}

const char *spg_mlp_last_error(const spg_mlp *net) { return net ? net->err.c_str() : g_mlpLastError.c_str(); }

int spg_mlp_get_params(spg_mlp *net, float *params)
{
    return guardedMlp(net, [&] {
        if (!net || !params) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        const auto flat = net->downloadFlat(net->P.p);
        std::memcpy(params, flat.data(), sizeof(float) * flat.size());
    });
}

int spg_mlp_set_params(spg_mlp *net, const float *params)
{
    return guardedMlp(net, [&] {
        if (!net || !params) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        net->setParams(params);
    });
}

int spg_mlp_get_normalization(spg_mlp *net, float *mean, float *std_dev)
{
    return guardedMlp(net, [&] {
        if (!net || !mean || !std_dev) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        std::copy(net->hMean.begin(), net->hMean.end(), mean);
        std::copy(net->hStd.begin(), net->hStd.end(), std_dev);
    });
}

int spg_mlp_set_history_base(spg_mlp *net, int32_t enabled)
{
    return guardedMlp(net, [&] {
        if (!net) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        net->historyBase = enabled ? 1 : 0;
    });
}

int spg_mlp_get_history_base(spg_mlp *net, int32_t *enabled)
{
    return guardedMlp(net, [&] {
        if (!net || !enabled) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        *enabled = net->historyBase;
    });
}

int spg_mlp_set_normalization(spg_mlp *net, const float *mean, const float *std_dev)
{
    return guardedMlp(net, [&] {
        if (!net || !mean || !std_dev) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        net->setNormalization(mean, std_dev);
    });
}

int spg_mlp_forward(spg_mlp *net, const float *features, int64_t n, float *out)
{
    return guardedMlp(net, [&] {
        if (!net || (n > 0 && (!features || !out))) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (n < 0) throw VbdError(SPG_VBD_ERR_INVALID, "negative row count");
        if (n == 0) return;
        requireFinite(features, n * kMlpFeatures, "features");
        DevBuf<float> f, y;
        f.alloc(n * kMlpFeatures);
        y.alloc(n * kMlpOut);
        hipCheck(hipMemcpyAsync(f.p, features, sizeof(float) * n * kMlpFeatures, hipMemcpyHostToDevice, net->stream),
                 "upload");
        net->infer(f.p, n, y.p, kMlpOut, net->stream);
        hipCheck(hipMemcpyAsync(out, y.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, net->stream), "download");
        hipCheck(hipStreamSynchronize(net->stream), "hipStreamSynchronize");
    });
}

int spg_mlp_loss_grad(spg_mlp *net, const float *features, const float *targets, int64_t n, double *loss, float *grad)
{
    return guardedMlp(net, [&] {
        if (!net || !features || !targets || !loss || !grad) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (n <= 0 || n > SPG_MLP_MAX_GRAD_ROWS) throw VbdError(SPG_VBD_ERR_INVALID, "row count out of range");
        requireFinite(features, n * kMlpFeatures, "features");
        requireFinite(targets, n * kMlpOut, "targets");
        const int64_t rows = roundUp(n, 64);
        net->reserveTrain(rows);
        hipStream_t s = net->stream;
        net->dataF.reserve(n * kMlpFeatures);
        net->dataT.reserve(n * kMlpOut);
        hipCheck(hipMemcpyAsync(net->dataF.p, features, sizeof(float) * n * kMlpFeatures, hipMemcpyHostToDevice, s),
                 "upload");
        hipCheck(hipMemcpyAsync(net->dataT.p, targets, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, s), "upload");
        hipCheck(launchPackInput(net->dataF.p, net->dataT.p, nullptr, net->mean.p, net->invStd.p, static_cast<int>(n),
                                 static_cast<int>(rows), net->X.p, net->XT.p, net->T.p, s),
                 "pack launch");
        hipCheck(hipMemsetAsync(net->loss.p, 0, sizeof(double), s), "memset");
        net->forward(rows, true, net->Yf.p, kMlpInferCols);
        net->backward(rows, static_cast<int>(n));
        double se = 0.0;
        hipCheck(hipMemcpyAsync(&se, net->loss.p, sizeof(double), hipMemcpyDeviceToHost, s), "download");
        const auto flat = net->downloadFlat(net->G.p);  // synchronises
        *loss = se / static_cast<double>(n);
        std::memcpy(grad, flat.data(), sizeof(float) * flat.size());
    });
}

int spg_mlp_train(spg_mlp *net, const float *features, const float *targets, int64_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, int32_t standardize, double *epoch_loss)
{
    return guardedMlp(net, [&] {
        if (!net || !features || !targets) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (n <= 0 || n > (int64_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_INVALID, "row count out of range");
        if (epochs < 0 || batch <= 0 || batch % 64 || !(lr > 0.f))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid epochs, batch (positive multiple of 64) or learning rate");
        requireFinite(features, n * kMlpFeatures, "features");
        requireFinite(targets, n * kMlpOut, "targets");
        if (standardize) {
            std::vector<double> sum(kMlpFeatures, 0.0), sq(kMlpFeatures, 0.0);
            for (int64_t i = 0; i < n; ++i)
                for (int k = 0; k < kMlpFeatures; ++k) {
                    const double v = features[i * kMlpFeatures + k];
                    sum[k] += v;
                    sq[k] += v * v;
                }
            std::vector<float> m(kMlpFeatures), sd(kMlpFeatures);
            for (int k = 0; k < kMlpFeatures; ++k) {
                const double mu = sum[k] / n;
                m[k] = static_cast<float>(mu);
                sd[k] = static_cast<float>(std::max(std::sqrt(std::max(sq[k] / n - mu * mu, 0.0)), 1e-6));
            }
            net->setNormalization(m.data(), sd.data());
        }
        hipStream_t s = net->stream;
        net->reserveTrain(batch);
        net->dataF.reserve(n * kMlpFeatures);
        net->dataT.reserve(n * kMlpOut);
        net->index.reserve(n);
        hipCheck(hipMemcpyAsync(net->dataF.p, features, sizeof(float) * n * kMlpFeatures, hipMemcpyHostToDevice, s),
                 "upload");
        hipCheck(hipMemcpyAsync(net->dataT.p, targets, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, s), "upload");
        const int64_t perEpoch = (n + batch - 1) / batch;
        const int64_t totalSteps = perEpoch * epochs;
        std::vector<int32_t> perm(n);
        uint64_t rs = seed;
        int64_t stepNo = 0;
        for (int32_t ep = 0; ep < epochs; ++ep) {
            for (int64_t i = 0; i < n; ++i) perm[i] = static_cast<int32_t>(i);
            for (int64_t i = n - 1; i > 0; --i) {  // Fisher-Yates
                const int64_t j = static_cast<int64_t>(uniform01(rs) * (i + 1));
                std::swap(perm[i], perm[std::min(j, i)]);
            }
            hipCheck(hipMemcpyAsync(net->index.p, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s), "upload");
            hipCheck(hipMemsetAsync(net->loss.p, 0, sizeof(double), s), "memset");
            for (int64_t b = 0; b < perEpoch; ++b, ++stepNo) {
                const int64_t r0 = b * batch;
                const int cnt = static_cast<int>(std::min<int64_t>(batch, n - r0));
                hipCheck(launchPackInput(net->dataF.p, net->dataT.p, net->index.p + r0, net->mean.p, net->invStd.p, cnt,
                                         batch, net->X.p, net->XT.p, net->T.p, s),
                         "pack launch");
                net->forward(batch, true, net->Yf.p, kMlpInferCols);
                net->backward(batch, cnt);
                const double frac = totalSteps > 1 ? static_cast<double>(stepNo) / (totalSteps - 1) : 0.0;
                const float lrNow = static_cast<float>(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * frac))));
                net->adamT += 1;
                hipCheck(launchAdam(net->P.p, net->G.p, net->M.p, net->V.p, net->total, lrNow, 0.9f, 0.999f, 1e-8f,
                                    net->adamT, s),
                         "adam launch");
                net->refreshWeights();
            }
            double se = 0.0;
            hipCheck(hipMemcpyAsync(&se, net->loss.p, sizeof(double), hipMemcpyDeviceToHost, s), "download");
            hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
            if (!std::isfinite(se)) throw VbdError(SPG_VBD_ERR_NONFINITE, "training diverged (non-finite loss)");
            if (epoch_loss) epoch_loss[ep] = se / static_cast<double>(n);
        }
        hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

}  // extern "C"
