// bt_model.h -- the degree model: a small MLP that maps the features of a
// far-field pair to the least interpolation degree that keeps the pair's
// error under a target.  Trained and evaluated with the HIP kernels of
// bt_learn.hip; host code here only stages data and runs the loop.
#pragma once

#include <string>
#include <vector>

#include "bt_learn.h"

namespace bt {

struct DegreeModel {
    std::vector<float> params = std::vector<float>(btl::MLP_PARAMS, 0.0f);
    float mean[btl::MLP_IN] = {};
    float scale[btl::MLP_IN] = {};  // all zero until set from training data
    // the range of every feature over the training data (lo > hi: unknown;
    // models saved before ranges were recorded have none)
    float lo[btl::MLP_IN], hi[btl::MLP_IN];
    int steps = 0;                  // Adam steps taken

    DegreeModel() { clear_range(); }
    void clear_range();
    bool has_range() const;
    void widen_range(int n, const float *x);  // the union with the range of x
    // 1 if a feature row lies outside the training range by more than 5% of
    // the range's width (the guard falls back to the fixed degree there)
    bool outside_range(const float *x) const;

    void init(unsigned long long seed);  // He-uniform weights, zero biases
    bool has_scaling() const;
    void set_scaling(int n, const float *x);  // mean and 1/std of the features

    std::string save(const std::string &path) const;
    std::string load(const std::string &path);

    // Batch-mean loss and its gradient (parameter layout) over n samples.
    std::string gradient(int n, const float *x, const int *labels, std::vector<float> &grad,
                         double &loss) const;
    // Adam over the data for `epochs` passes of shuffled batches; epoch_loss
    // (optional) gets the mean loss of each epoch.
    std::string train(int n, const float *x, const int *labels, int epochs, int batch, float lr,
                      unsigned long long seed, std::vector<float> *epoch_loss);
    // Degrees (and optionally logits, 10 per sample) for n feature rows.
    std::string predict(int n, const float *x, float delta, int *degree, float *logits) const;
};

}  // namespace bt
