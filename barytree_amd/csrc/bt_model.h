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
    int steps = 0;                  // Adam steps taken

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
