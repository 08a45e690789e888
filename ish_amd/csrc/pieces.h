// Host-side gather of a database upload: the record bytes of a batch as
// pieces in separate host buffers (the file reader's segments), uploaded
// one after the other into one device buffer.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/ish_hip.h"

namespace ish {

struct HostPiece {
    const uint8_t *p;
    uint64_t n;
};

// ish_hip_db_load_ascii over the concatenation of pieces[0..np): offsets
// (n + 1 entries, offsets[0] = 0) index that concatenation, whose length
// offsets[n] must equal the pieces' total. With slot 0 or 1 the database
// borrows the context's one-shot buffers of that slot (grow-only, never
// freed mid-run); it must be done with (scored and fetched) before the
// slot is loaded again.
int32_t db_load_ascii_pieces(ish_hip_ctx *ctx, const HostPiece *pieces, size_t np,
                             const uint64_t *offsets, int64_t n, int32_t matrix_kind,
                             ish_hip_db **out, int slot = -1);

// Compute units of the context's device, and the device.
int32_t ctx_cus(const ish_hip_ctx *ctx);
int32_t ctx_device(const ish_hip_ctx *ctx);

// Bracket a one-shot call (slot loads, db_score_once_async,
// db_fetch_after): begin creates the context's copy stream, end waits for
// it and destroys it, so that the context holds one hardware queue between
// calls.
int32_t one_shot_begin(ish_hip_ctx *ctx);
void one_shot_end(ish_hip_ctx *ctx);

// One-shot passes: db_score_once_async enqueues a pass (no graph capture)
// behind whatever the context's stream holds and marks its end with the
// event of `slot`; db_fetch_after copies its results on the copy stream
// once that event fires, so a pass queued after it keeps the device busy
// meanwhile. Returns when the results are in the host arrays.
int32_t db_score_once_async(ish_hip_ctx *ctx, ish_hip_db *db, int slot);
int32_t db_fetch_after(ish_hip_ctx *ctx, ish_hip_db *db, int slot, int32_t *score,
                       int32_t *end_ref, int32_t *end_q, uint8_t *status);

} // namespace ish
