// tables.h -- host-side construction of every lookup table the move
// generator uses (the work of MoveGeneratorBitboard::init,
// MoveGeneratorBitboard.h:982-1080, and Magics.cpp:27-150).
#pragma once
#include <cstddef>
#include <vector>
#include "movegen.h"

namespace pg {

struct HostTables {
    u64 knight[64], king[64], rookEmpty[64], bishopEmpty[64];
    u64 between[64 * 64], line[64 * 64];
    u64 rookMask[64], bishopMask[64];
    u64 rookMagic[64], bishopMagic[64];
    std::vector<u64> rookPlain, bishopPlain;      // 64*4096, 64*512
    FancyEntry rookFancy[64], bishopFancy[64];
    std::vector<u64> fancy;                       // 97264
    std::vector<u8> byteIdx;                      // 97264
    std::vector<u64> byteRook, byteBishop;        // 4900, 1428
    u64 zobLo[ZOB_WORDS], zobHi[ZOB_WORDS];
    u64 magicSeed = 0;

    // Builds everything; returns 0, or -1 with a message in err.
    int build(u64 seed, char *err, int errlen);
    Tables view() const;
};

// Shared, lazily built host tables (seed fixed so builds are reproducible).
// Returns null, with a message in *err, if construction failed.
const HostTables *host_tables(const char **err = nullptr);

// Serialized byte layout used for the single device allocation.
struct TableLayout {
    size_t off[20];
    size_t total;
};
TableLayout table_layout();
void pack_tables(const HostTables &h, std::vector<u8> &blob);
Tables tables_at(const u8 *base);  // pointers into a packed blob

}  // namespace pg
