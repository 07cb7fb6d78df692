// Host-side sampler tables of the synthetic corpus generator.
#pragma once
#include <stdint.h>

#include <vector>

namespace w2v {

struct SynthTables {
  std::vector<uint64_t> gthr;    // [V] global thresholds, index = word id
  std::vector<int32_t> members;  // [V-1] word ids grouped by cluster
  std::vector<int64_t> cstart;   // [C+1]
  std::vector<uint64_t> cthr;    // [V-1]
  std::vector<uint64_t> lthr;    // [max+1]
  uint64_t stay_thr = 0, in_thr = 0;
};

}  // namespace w2v
