// Bounds of the row-major numeric assembly (k_asm_rows): one single-wave
// workgroup keeps one entity's CSR rows, its pair list, its element list and
// the column of every (element, local entity) in LDS.  Plans with an entity
// beyond them use the two-phase kernels instead.  Rows of at most
// kRowAsmShortLen columns use a variant with a quarter less LDS per workgroup
// (more resident workgroups).
#pragma once

namespace pg {

constexpr int kRowAsmMaxLen = 192;   // columns of an entity's rows
constexpr int kRowAsmShortLen = 96;  // the short-row variant
constexpr int kRowAsmMaxInc = 24;    // elements containing one entity
constexpr int kRowAsmMaxPairs = 96;  // entities coupled to one entity

}  // namespace pg
