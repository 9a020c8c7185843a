// host_batch.h -- the store + repair of one batch of host-buffer updates, as
// the request server (server.hip) drives it.
#pragma once
#include <stdint.h>

#include <functional>
#include <string>

#include "../../include/risgraph_amd.h"

namespace rg {

// Applies updates [0, n) (edges e, kind +1 add / -1 del, txn_of[i] = the
// transaction of update i, for the log) as one device batch on the labels
// L[0..nL): logged first when the graph has a WAL open, then stored,
// classified and repaired.  flags[i] (may be null) = update i waited for the
// repair (unsafe); on_safe (may be empty) is called with the flags as soon
// as the classification is on the host, while the repair still runs.
// *v_before / *v_after = the graph's version before and after the batch.
// Returns an RG_* status (err gets the message).
int host_batch(rg_graph *g, rg_labels **L, int nL, const rg_edge *e, const int8_t *kind, const uint32_t *txn_of,
               uint64_t n, int directed, uint8_t *flags, const std::function<void(const uint8_t *)> &on_safe,
               uint64_t *v_before, uint64_t *v_after, std::string *err);

}  // namespace rg
