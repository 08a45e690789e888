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
// offsets[n] must equal the pieces' total.
int32_t db_load_ascii_pieces(ish_hip_ctx *ctx, const HostPiece *pieces, size_t np,
                             const uint64_t *offsets, int64_t n, int32_t matrix_kind,
                             ish_hip_db **out);

// ish_hip_align_db_async without graph capture, for databases scored once.
int32_t db_score_once_async(ish_hip_ctx *ctx, ish_hip_db *db);

} // namespace ish
