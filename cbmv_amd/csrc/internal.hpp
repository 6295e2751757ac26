// Declarations shared between translation units of libcbmv_hip.so.
#pragma once

#include <vector>

#include "forest.hpp"
#include "ops.hpp"

namespace cbmv {

ForestView forest_view(const cbmv_forest* f);
const std::vector<GpuNode>& forest_host_nodes(const cbmv_forest* f);
const std::vector<double>& forest_host_leaf0(const cbmv_forest* f);
const std::vector<int32_t>& forest_host_roots(const cbmv_forest* f);
int forest_n_trees(const cbmv_forest* f);
int forest_n_features(const cbmv_forest* f);

}  // namespace cbmv
