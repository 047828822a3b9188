// Learned ACA rank predictor: block features, network state on the device,
// training and inference. Used by the assembly only to presize the ACA
// workspace of each admissible block and to order blocks by expected work;
// ACA's own stopping test decides the rank.
#pragma once

#include "common.hpp"
#include "mlp.hpp"
#include "tree.hpp"

#include <vector>

struct htg_predictor {
    htg::DevArray<float> params, m, v; // weights, Adam moments
    htg::DevArray<float> mean, stdv;   // feature normalisation
    float h_mean[htg::MLP_IN] = {0}, h_std[htg::MLP_IN] = {1, 1, 1, 1, 1, 1, 1, 1};
    int step = 0; // Adam steps taken
};

namespace htg {

// The 8 features of an admissible block (t, s) at (eta, epsilon):
// ln m, ln n, ln d, ln diam_t, ln diam_s, eta, ln epsilon, ln gap, with d the
// distance between the cluster centres, diam = 2 x radius, and gap =
// max(d - rad_t - rad_s, 1e-6 (rad_t + rad_s)) the clearance between the
// clusters' bounding balls.
void block_features(const ClusterNode &t, const ClusterNode &s, double eta, double epsilon, float *f);

// Rank-like training target of an ACA outcome: the number of cross terms the
// block's ACA run needed, i.e. its rank, or floor(mn / (m + n)) when it
// stopped as not compressible (rank -1).
int aca_terms_needed(int m, int n, int rank);

// Expected terms for each feature row (host arrays), on the device.
void predictor_predict(const htg_predictor &p, const float *feat, int n, float *out_terms);

} // namespace htg
