// bt_model.cpp -- host side of the degree model (bt_model.h).
#include "bt_model.h"
#include "bt_devmem.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>

namespace bt {

namespace {

// file layouts: 01 without the training range, 02 with it
constexpr char MAGIC1[8] = {'B', 'T', 'D', 'M', 'L', 'P', '0', '1'};
constexpr char MAGIC[8] = {'B', 'T', 'D', 'M', 'L', 'P', '0', '2'};

// Device buffers owned by one call.
struct DevBufs {
    std::vector<void *> p;
    // the call's work on the device is finished before the buffers go back
    // to the device-memory cache
    ~DevBufs()
    {
        if (p.empty()) return;
        (void)hipDeviceSynchronize();
        for (void *q : p) btm::release(q);
    }
    template <class T> T *get(size_t n)
    {
        void *q = btm::alloc(std::max<size_t>(n, 1) * sizeof(T));
        if (!q) return nullptr;
        p.push_back(q);
        return static_cast<T *>(q);
    }
};

#define CHK(expr)                                                                     \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
    } while (0)

bool alloc_train(DevBufs &db, int b, btl::TrainBufs &tb)
{
    const size_t H = (size_t)btl::MLP_H * b;
    tb.a0 = db.get<float>((size_t)btl::MLP_IN * b);
    tb.a1 = db.get<float>(H); tb.a2 = db.get<float>(H); tb.a3 = db.get<float>(H);
    tb.d1 = db.get<float>(H); tb.d2 = db.get<float>(H); tb.d3 = db.get<float>(H);
    tb.dz = db.get<float>((size_t)btl::MLP_OUT * b);
    tb.loss = db.get<float>(b);
    return tb.a0 && tb.a1 && tb.a2 && tb.a3 && tb.d1 && tb.d2 && tb.d3 && tb.dz && tb.loss;
}

}  // namespace

void DegreeModel::init(unsigned long long seed)
{
    std::mt19937_64 rng(seed);
    std::fill(params.begin(), params.end(), 0.0f);
    auto fill = [&](int off, int rows, int cols) {
        const double a = std::sqrt(6.0 / cols);
        std::uniform_real_distribution<double> u(-a, a);
        for (int i = 0; i < rows * cols; i++) params[off + i] = (float)u(rng);
    };
    fill(btl::OFF_W1, btl::MLP_H, btl::MLP_IN);
    fill(btl::OFF_W2, btl::MLP_H, btl::MLP_H);
    fill(btl::OFF_W3, btl::MLP_H, btl::MLP_H);
    fill(btl::OFF_W4, btl::MLP_OUT, btl::MLP_H);
    steps = 0;
}

void DegreeModel::clear_range()
{
    for (int f = 0; f < btl::MLP_IN; f++) { lo[f] = 1.0f; hi[f] = -1.0f; }
}

bool DegreeModel::has_range() const
{
    for (int f = 0; f < btl::MLP_IN; f++)
        if (!(lo[f] <= hi[f])) return false;
    return true;
}

void DegreeModel::widen_range(int n, const float *x)
{
    for (int f = 0; f < btl::MLP_IN; f++) {
        float a = lo[f], b = hi[f];
        bool have = a <= b;
        for (int i = 0; i < n; i++) {
            const float v = x[(size_t)i * btl::MLP_IN + f];
            if (!std::isfinite(v)) continue;
            if (!have) { a = b = v; have = true; continue; }
            a = std::min(a, v);
            b = std::max(b, v);
        }
        if (have) { lo[f] = a; hi[f] = b; }
    }
}

bool DegreeModel::outside_range(const float *x) const
{
    if (!has_range()) return false;
    for (int f = 0; f < btl::MLP_IN; f++) {
        const float m = 0.05f * (hi[f] - lo[f]);
        if (!(x[f] >= lo[f] - m && x[f] <= hi[f] + m)) return true;  // NaN counts as outside
    }
    return false;
}

bool DegreeModel::has_scaling() const
{
    for (float s : scale)
        if (s != 0.0f) return true;
    return false;
}

void DegreeModel::set_scaling(int n, const float *x)
{
    for (int f = 0; f < btl::MLP_IN; f++) {
        double m = 0.0, v = 0.0;
        for (int i = 0; i < n; i++) m += x[(size_t)i * btl::MLP_IN + f];
        m = n ? m / n : 0.0;
        for (int i = 0; i < n; i++) {
            const double d = x[(size_t)i * btl::MLP_IN + f] - m;
            v += d * d;
        }
        const double sd = n ? std::sqrt(v / n) : 1.0;
        mean[f] = (float)m;
        scale[f] = (float)(1.0 / std::max(sd, 1e-6));
    }
}

std::string DegreeModel::save(const std::string &path) const
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return "cannot open " + path + " for writing";
    const int hdr[4] = {btl::MLP_IN, btl::MLP_H, btl::MLP_OUT, btl::MLP_PARAMS};
    bool ok = std::fwrite(MAGIC, 1, 8, f) == 8 && std::fwrite(hdr, sizeof(int), 4, f) == 4 &&
              std::fwrite(&steps, sizeof(int), 1, f) == 1 &&
              std::fwrite(mean, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
              std::fwrite(scale, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
              std::fwrite(lo, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
              std::fwrite(hi, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
              std::fwrite(params.data(), sizeof(float), params.size(), f) == params.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? "" : "write to " + path + " failed";
}

std::string DegreeModel::load(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return "cannot open " + path;
    char mg[8];
    int hdr[4];
    DegreeModel m;
    bool ok = std::fread(mg, 1, 8, f) == 8 && (std::memcmp(mg, MAGIC, 8) == 0 || std::memcmp(mg, MAGIC1, 8) == 0);
    const bool ranged = ok && std::memcmp(mg, MAGIC, 8) == 0;
    ok = ok && std::fread(hdr, sizeof(int), 4, f) == 4 && hdr[0] == btl::MLP_IN &&
         hdr[1] == btl::MLP_H && hdr[2] == btl::MLP_OUT && hdr[3] == btl::MLP_PARAMS &&
         std::fread(&m.steps, sizeof(int), 1, f) == 1 &&
         std::fread(m.mean, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
         std::fread(m.scale, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
         (!ranged || (std::fread(m.lo, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN &&
                      std::fread(m.hi, sizeof(float), btl::MLP_IN, f) == (size_t)btl::MLP_IN)) &&
         std::fread(m.params.data(), sizeof(float), m.params.size(), f) == m.params.size();
    std::fclose(f);
    if (!ok) return path + " is not a degree model file of this layout";
    *this = m;
    return "";
}

std::string DegreeModel::gradient(int n, const float *x, const int *labels, std::vector<float> &grad,
                                  double &loss) const
{
    if (n <= 0) return "no samples";
    if (!has_scaling()) return "model has no feature scaling (train it or set it first)";
    DevBufs db;
    float *dp = db.get<float>(btl::MLP_PARAMS), *dmu = db.get<float>(btl::MLP_IN),
          *dsc = db.get<float>(btl::MLP_IN), *dx = db.get<float>((size_t)n * btl::MLP_IN),
          *dg = db.get<float>(btl::MLP_PARAMS);
    int *dl = db.get<int>(n);
    btl::TrainBufs tb;
    if (!dp || !dmu || !dsc || !dx || !dg || !dl || !alloc_train(db, n, tb)) return "hipMalloc failed";
    CHK(hipMemcpy(dp, params.data(), btl::MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dmu, mean, sizeof(mean), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dsc, scale, sizeof(scale), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dx, x, (size_t)n * btl::MLP_IN * sizeof(float), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dl, labels, n * sizeof(int), hipMemcpyHostToDevice));
    CHK(btl::launch_mlp_fwdbwd(dp, dmu, dsc, dx, dl, nullptr, n, tb, nullptr));
    CHK(btl::launch_mlp_wgrad(tb, n, dg, nullptr));
    grad.assign(btl::MLP_PARAMS, 0.0f);
    std::vector<float> ls(n);
    CHK(hipMemcpy(grad.data(), dg, btl::MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    CHK(hipMemcpy(ls.data(), tb.loss, n * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0.0;
    for (float v : ls) s += v;
    loss = s / n;
    return "";
}

std::string DegreeModel::train(int n, const float *x, const int *labels, int epochs, int batch, float lr,
                               unsigned long long seed, std::vector<float> *epoch_loss)
{
    if (n <= 0 || epochs <= 0 || batch <= 0) return "bad training arguments";
    if (!has_scaling()) set_scaling(n, x);
    widen_range(n, x);
    batch = std::min(batch, n);
    DevBufs db;
    float *dp = db.get<float>(btl::MLP_PARAMS), *dmu = db.get<float>(btl::MLP_IN),
          *dsc = db.get<float>(btl::MLP_IN), *dx = db.get<float>((size_t)n * btl::MLP_IN),
          *dg = db.get<float>(btl::MLP_PARAMS), *dm = db.get<float>(btl::MLP_PARAMS),
          *dv = db.get<float>(btl::MLP_PARAMS);
    int *dl = db.get<int>(n), *didx = db.get<int>(n);
    btl::TrainBufs tb;
    if (!dp || !dmu || !dsc || !dx || !dg || !dm || !dv || !dl || !didx || !alloc_train(db, batch, tb))
        return "hipMalloc failed";
    CHK(hipMemcpy(dp, params.data(), btl::MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dmu, mean, sizeof(mean), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dsc, scale, sizeof(scale), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dx, x, (size_t)n * btl::MLP_IN * sizeof(float), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dl, labels, n * sizeof(int), hipMemcpyHostToDevice));
    CHK(hipMemset(dm, 0, btl::MLP_PARAMS * sizeof(float)));
    CHK(hipMemset(dv, 0, btl::MLP_PARAMS * sizeof(float)));
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<float> ls(batch);
    if (epoch_loss) epoch_loss->clear();
    const int nsteps = n / batch;  // full batches only
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        CHK(hipMemcpy(didx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
        double es = 0.0;
        for (int k = 0; k < std::max(1, nsteps); k++) {
            CHK(btl::launch_mlp_fwdbwd(dp, dmu, dsc, dx, dl, didx + (size_t)k * batch, batch, tb, nullptr));
            CHK(btl::launch_mlp_wgrad(tb, batch, dg, nullptr));
            CHK(btl::launch_adam(dp, dg, dm, dv, btl::MLP_PARAMS, lr, 0.9f, 0.999f, 1e-8f, ++steps, nullptr));
            CHK(hipMemcpy(ls.data(), tb.loss, batch * sizeof(float), hipMemcpyDeviceToHost));
            double s = 0.0;
            for (float v : ls) s += v;
            es += s / batch;
        }
        if (epoch_loss) epoch_loss->push_back((float)(es / std::max(1, nsteps)));
    }
    CHK(hipMemcpy(params.data(), dp, btl::MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    return "";
}

std::string DegreeModel::predict(int n, const float *x, float delta, int *degree, float *logits) const
{
    if (n < 0) return "bad sample count";
    if (!(delta > 0.0f && delta < 1.0f)) return "delta must lie in (0, 1)";
    if (n == 0) return "";
    if (!has_scaling()) return "model has no feature scaling (train it or set it first)";
    // in chunks, so the activation buffers stay small
    const int chunk = std::min(n, 1 << 18);
    DevBufs db;
    float *dp = db.get<float>(btl::MLP_PARAMS), *dmu = db.get<float>(btl::MLP_IN),
          *dsc = db.get<float>(btl::MLP_IN), *dx = db.get<float>((size_t)chunk * btl::MLP_IN);
    int *dd = db.get<int>(chunk);
    float *dz = logits ? db.get<float>((size_t)chunk * btl::MLP_OUT) : nullptr;
    btl::TrainBufs tb;
    if (!dp || !dmu || !dsc || !dx || !dd || (logits && !dz) || !alloc_train(db, chunk, tb))
        return "hipMalloc failed";
    CHK(hipMemcpy(dp, params.data(), btl::MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dmu, mean, sizeof(mean), hipMemcpyHostToDevice));
    CHK(hipMemcpy(dsc, scale, sizeof(scale), hipMemcpyHostToDevice));
    for (int o = 0; o < n; o += chunk) {
        const int m = std::min(chunk, n - o);
        CHK(hipMemcpy(dx, x + (size_t)o * btl::MLP_IN, (size_t)m * btl::MLP_IN * sizeof(float),
                      hipMemcpyHostToDevice));
        CHK(btl::launch_mlp_infer(dp, dmu, dsc, m, dx, delta, dz, dd, tb, nullptr));
        if (degree) CHK(hipMemcpy(degree + o, dd, m * sizeof(int), hipMemcpyDeviceToHost));
        if (logits)
            CHK(hipMemcpy(logits + (size_t)o * btl::MLP_OUT, dz, (size_t)m * btl::MLP_OUT * sizeof(float),
                          hipMemcpyDeviceToHost));
    }
    return "";
}

}  // namespace bt
