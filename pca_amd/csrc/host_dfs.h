// host_dfs.h -- CPU prefix producer: the branch-mode IDDFS of
// gpu/hybrid.cu:524-588 (tools/enumerate.c:405-461) down to the split depth,
// with exact OEP keys and incremental per-test state in place of the
// reference's VM replay.  Emits surviving prefixes in DFS order.
#pragma once
#include <stdint.h>
#include <functional>
#include <vector>

#include "host_core.h"

namespace pca {

// Open-addressing set of fixed-length byte keys (the per-test OEP state
// bytes); exact: equal hashes are confirmed byte-for-byte.
class ExactKeySet {
public:
    explicit ExactKeySet(int klen = 1) : klen_(klen) {}
    void reset(int klen);
    void clear();
    // 1 = inserted (first visitor), 0 = already present
    int insert(const uint8_t *key);
    size_t size() const { return n_; }

private:
    void grow();
    int klen_;
    std::vector<uint64_t> hash_;
    std::vector<uint32_t> slot_;
    std::vector<uint8_t> arena_;
    size_t n_ = 0;
};

struct DfsStats {
    uint64_t nodes = 0;     // tried counter of the prefix levels
    uint64_t pruned = 0;    // OEP prunes
    uint64_t emitted = 0;   // prefixes emitted at the split
};

// Runs the depth-D tree down to split s.  `emit` receives each surviving
// length-s prefix in DFS order.  Returning false from emit stops the walk.
void dfs_prefixes(const pca_task_t *task, int D, int s,
                  const std::function<bool(const pca_prefix_t &)> &emit, DfsStats *stats);

}  // namespace pca
