// Lending of a pool's idle worker contexts to the HM subtree solves of a job
// running on another worker (library-internal: pool.cpp and the context
// code; not part of include/heipa.h).  borrow returns up to `max` idle
// contexts in `out` and marks them lent, so no queued job starts on them
// until give_back returns them.
#pragma once

#include "../../../include/heipa.h"

#include <cstddef>

extern "C" {
typedef size_t (*heipa_borrow_fn)(void *user, heipa_ctx **out, size_t max);
typedef void (*heipa_give_back_fn)(void *user, heipa_ctx *const *ctxs, size_t n);
int heipa_ctx_set_lender_internal(heipa_ctx *ctx, heipa_borrow_fn borrow, heipa_give_back_fn give_back, void *user);
}
