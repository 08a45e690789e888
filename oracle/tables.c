/*
 * Scoring matrices and alphabet encodings used by the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The values are data that must be identical to the reference's tables:
 *   BLOSUM62 (24x24, order ARNDCQEGHILKMFPSTWYVBZX*):
 *       ishlib/matcher/alignment/scoring_matrix.mojo:98-124
 *   ACTGN / ACTGN0 (5x5):            scoring_matrix.mojo:156-172
 *   AA_TO_NUM / NUM_TO_AA:           scoring_matrix.mojo:9-35
 *   NT_TO_NUM / NUM_TO_NT:           scoring_matrix.mojo:126-149
 *   ascii default (256x256, +2/-2):  scoring_matrix.mojo:360-391
 * The ascii->code tables are generated here from their letter rules rather
 * than listed byte by byte: amino acids map to their index in
 * "ARNDCQEGHILKMFPSTWYVBZX" (either case), everything else to 23 ('*');
 * nucleotides map A0 C1 G2 T3 U3 (either case), everything else to 4 (N).
 */
#include <string.h>
#include "oracle.h"

static const int8_t BLOSUM62_T[24 * 24] = {
/*        A   R   N   D   C   Q   E   G   H   I   L   K   M   F   P   S   T   W   Y   V   B   Z   X   * */
/* A */   4, -1, -2, -2,  0, -1, -1,  0, -2, -1, -1, -1, -1, -2, -1,  1,  0, -3, -2,  0, -2, -1,  0, -4,
/* R */  -1,  5,  0, -2, -3,  1,  0, -2,  0, -3, -2,  2, -1, -3, -2, -1, -1, -3, -2, -3, -1,  0, -1, -4,
/* N */  -2,  0,  6,  1, -3,  0,  0,  0,  1, -3, -3,  0, -2, -3, -2,  1,  0, -4, -2, -3,  3,  0, -1, -4,
/* D */  -2, -2,  1,  6, -3,  0,  2, -1, -1, -3, -4, -1, -3, -3, -1,  0, -1, -4, -3, -3,  4,  1, -1, -4,
/* C */   0, -3, -3, -3,  9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1, -3, -3, -2, -4,
/* Q */  -1,  1,  0,  0, -3,  5,  2, -2,  0, -3, -2,  1,  0, -3, -1,  0, -1, -2, -1, -2,  0,  3, -1, -4,
/* E */  -1,  0,  0,  2, -4,  2,  5, -2,  0, -3, -3,  1, -2, -3, -1,  0, -1, -3, -2, -2,  1,  4, -1, -4,
/* G */   0, -2,  0, -1, -3, -2, -2,  6, -2, -4, -4, -2, -3, -3, -2,  0, -2, -2, -3, -3, -1, -2, -1, -4,
/* H */  -2,  0,  1, -1, -3,  0,  0, -2,  8, -3, -3, -1, -2, -1, -2, -1, -2, -2,  2, -3,  0,  0, -1, -4,
/* I */  -1, -3, -3, -3, -1, -3, -3, -4, -3,  4,  2, -3,  1,  0, -3, -2, -1, -3, -1,  3, -3, -3, -1, -4,
/* L */  -1, -2, -3, -4, -1, -2, -3, -4, -3,  2,  4, -2,  2,  0, -3, -2, -1, -2, -1,  1, -4, -3, -1, -4,
/* K */  -1,  2,  0, -1, -3,  1,  1, -2, -1, -3, -2,  5, -1, -3, -1,  0, -1, -3, -2, -2,  0,  1, -1, -4,
/* M */  -1, -1, -2, -3, -1,  0, -2, -3, -2,  1,  2, -1,  5,  0, -2, -1, -1, -1, -1,  1, -3, -1, -1, -4,
/* F */  -2, -3, -3, -3, -2, -3, -3, -3, -1,  0,  0, -3,  0,  6, -4, -2, -2,  1,  3, -1, -3, -3, -1, -4,
/* P */  -1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4,  7, -1, -1, -4, -3, -2, -2, -1, -2, -4,
/* S */   1, -1,  1,  0, -1,  0,  0,  0, -1, -2, -2,  0, -1, -2, -1,  4,  1, -3, -2, -2,  0,  0,  0, -4,
/* T */   0, -1,  0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1,  1,  5, -2, -2,  0, -1, -1,  0, -4,
/* W */  -3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1,  1, -4, -3, -2, 11,  2, -3, -4, -3, -2, -4,
/* Y */  -2, -2, -2, -3, -2, -1, -2, -3,  2, -1, -1, -2, -1,  3, -3, -2, -2,  2,  7, -1, -3, -2, -1, -4,
/* V */   0, -3, -3, -3, -1, -2, -2, -3, -3,  3,  1, -2,  1, -1, -2, -2,  0, -3, -1,  4, -3, -2, -1, -4,
/* B */  -2, -1,  3,  4, -3,  0,  1, -1,  0, -3, -4,  0, -3, -3, -2,  0, -1, -4, -3, -3,  4,  1, -1, -4,
/* Z */  -1,  0,  0,  1, -3,  3,  4, -2,  0, -3, -3,  1, -1, -3, -1,  0, -1, -3, -2, -2,  1,  4, -1, -4,
/* X */   0, -1, -1, -1, -2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -2,  0,  0, -2, -1, -1, -1, -1, -1, -4,
/* * */  -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4,  1,
};

/* +2 on the ACGT diagonal, -2 off it, and an N row/column of +2 (actgn) or
 * 0 (actgn0). */
static const int8_t ACTGN_T[25] = {
     2, -2, -2, -2, 2,
    -2,  2, -2, -2, 2,
    -2, -2,  2, -2, 2,
    -2, -2, -2,  2, 2,
     2,  2,  2,  2, 2,
};
static const int8_t ACTGN0_T[25] = {
     2, -2, -2, -2, 0,
    -2,  2, -2, -2, 0,
    -2, -2,  2, -2, 0,
    -2, -2, -2,  2, 0,
     0,  0,  0,  0, 0,
};

static const char AA_LETTERS[] = "ARNDCQEGHILKMFPSTWYVBZX*";
static const char NT_LETTERS[] = "ACGTN";

int32_t oracle_alphabet_size(int32_t kind) {
    switch (kind) {
    case 0: return 256;
    case 1: case 2: return 5;
    case 3: return 24;
    default: return -1;
    }
}

int32_t oracle_matrix(int32_t kind, int8_t *out) {
    switch (kind) {
    case 0:
        for (int a = 0; a < 256; ++a)
            for (int b = 0; b < 256; ++b) out[a * 256 + b] = (a == b) ? 2 : -2;
        return 256;
    case 1: memcpy(out, ACTGN_T, 25); return 5;
    case 2: memcpy(out, ACTGN0_T, 25); return 5;
    case 3: memcpy(out, BLOSUM62_T, 576); return 24;
    default: return -1;
    }
}

int32_t oracle_encoding(int32_t kind, uint8_t *out) {
    switch (kind) {
    case 0:
        for (int c = 0; c < 256; ++c) out[c] = (uint8_t)c;
        return 0;
    case 1: case 2:
        for (int c = 0; c < 256; ++c) out[c] = 4;
        out['A'] = out['a'] = 0;
        out['C'] = out['c'] = 1;
        out['G'] = out['g'] = 2;
        out['T'] = out['t'] = 3;
        out['U'] = out['u'] = 3;
        return 0;
    case 3:
        for (int c = 0; c < 256; ++c) out[c] = 23;
        for (int k = 0; k < 23; ++k) {
            out[(unsigned char)AA_LETTERS[k]] = (uint8_t)k;
            out[(unsigned char)(AA_LETTERS[k] + ('a' - 'A'))] = (uint8_t)k;
        }
        return 0;
    default: return -1;
    }
}

int32_t oracle_decoding(int32_t kind, uint8_t *out) {
    switch (kind) {
    case 0:
        for (int c = 0; c < 256; ++c) out[c] = (uint8_t)c;
        return 256;
    case 1: case 2: memcpy(out, NT_LETTERS, 5); return 5;
    case 3: memcpy(out, AA_LETTERS, 24); return 24;
    default: return -1;
    }
}

int64_t oracle_encode_and_score(int32_t kind, const uint8_t *ascii, int64_t n,
                                uint8_t *out) {
    static _Thread_local int8_t m[65536];
    uint8_t enc[256];
    int32_t A = oracle_matrix(kind, m);
    if (A < 0 || oracle_encoding(kind, enc) != 0) return INT64_MIN;
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint8_t c = enc[ascii[i]];
        out[i] = c;
        s += m[c * A + c];
    }
    return s;
}
