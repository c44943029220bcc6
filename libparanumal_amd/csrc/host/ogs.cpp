// Single-rank gather-scatter setup and device application.
#include "ogs.hpp"

#include <algorithm>

namespace libp {

randStream_t::randStream_t(unsigned seed) {
  int32_t r[34];
  r[0] = (int32_t)(seed == 0 ? 1 : seed);
  for (int i = 1; i < 31; ++i) {
    // 16807 * r mod (2^31 - 1) without overflow (Schrage)
    const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
    int64_t word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r[i] = (int32_t)word;
  }
  for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
  for (int i = 0; i < 34; ++i) ring[i] = (uint32_t)r[i];
  pos = 0;
  for (int i = 34; i < 344; ++i) next();
}

int randStream_t::next() {
  const uint32_t v = ring[(pos + 3) % 34] + ring[(pos + 31) % 34];
  ring[pos] = v;
  pos = (pos + 1) % 34;
  return (int)(v >> 1);
}

namespace ogs {

void ogs_t::Setup(dlong _N, memory<hlong> ids, bool _unique, randStream_t& rng, platform_t& _platform) {
  platform = _platform;
  N = _N;
  unique = _unique;
  LIBP_ABORT("ogs_t::Setup: ids shorter than N", (dlong)ids.length() < N);

  // Compressed list of non-zero ids in local order.
  std::vector<dlong> local;
  local.reserve(N);
  hlong maxId = 0;
  for (dlong n = 0; n < N; ++n)
    if (ids[n] != 0) {
      local.push_back(n);
      maxId = std::max(maxId, ids[n] < 0 ? -ids[n] : ids[n]);
    }
  const dlong Nids = (dlong)local.size();
  LIBP_ABORT("ogs_t::Setup: id range too large for the single-rank map", maxId > (hlong)2 * N + 2);

  // Owner of each group (FindSharedNodes, ogsSetup.cpp:241-275): the records,
  // in local order, are sorted by |id| with an unstable sort, and within each
  // group the copy at position rand() % size becomes the owner. The sort only
  // ever compares keys, so sorting (key, index) pairs reproduces the
  // reference's permutation exactly.
  std::vector<char> isOwner(Nids, 1);
  if (unique) {
    struct Rec {
      hlong key;
      dlong c;
    };
    std::vector<Rec> rec(Nids);
    for (dlong c = 0; c < Nids; ++c) {
      const hlong id = ids[local[c]];
      rec[c] = Rec{id < 0 ? -id : id, c};
    }
    std::sort(rec.begin(), rec.end(), [](const Rec& a, const Rec& b) { return a.key < b.key; });
    std::fill(isOwner.begin(), isOwner.end(), 0);
    dlong start = 0;
    for (dlong n = 0; n < Nids; ++n) {
      if (n == Nids - 1 || rec[n].key != rec[n + 1].key) {
        const dlong end = n + 1;
        const int m = rng.next() % (end - start);
        isOwner[rec[start + m].c] = 1;
        start = end;
      }
    }
  }

  // Rows in order of first encounter in local order (ogsSetup.cpp:409-433).
  std::vector<dlong> rowOf((size_t)maxId + 1, -1);
  std::vector<dlong> row(Nids);
  dlong Nrows = 0;
  for (dlong c = 0; c < Nids; ++c) {
    const hlong id = ids[local[c]];
    const hlong a = id < 0 ? -id : id;
    if (rowOf[a] < 0) rowOf[a] = Nrows++;
    row[c] = rowOf[a];
  }
  NlocalT = NlocalP = Nrows;
  Ngather = Nrows;
  NgatherGlobal = Nrows;

  // CSR (LocalSignedSetup, ogsSetup.cpp:569-683): columns in ascending local order.
  rowStartsT.malloc(Nrows + 1, 0);
  rowStartsN.malloc(Nrows + 1, 0);
  for (dlong c = 0; c < Nids; ++c) {
    rowStartsT[row[c] + 1]++;
    if (isOwner[c]) rowStartsN[row[c] + 1]++;
  }
  for (dlong r = 0; r < Nrows; ++r) {
    rowStartsT[r + 1] += rowStartsT[r];
    rowStartsN[r + 1] += rowStartsN[r];
  }
  colIdsT.malloc(rowStartsT[Nrows]);
  colIdsN.malloc(rowStartsN[Nrows]);
  std::vector<dlong> fillT(rowStartsT.ptr(), rowStartsT.ptr() + Nrows), fillN(rowStartsN.ptr(), rowStartsN.ptr() + Nrows);
  for (dlong c = 0; c < Nids; ++c) {
    colIdsT[fillT[row[c]]++] = local[c];
    if (isOwner[c]) colIdsN[fillN[row[c]]++] = local[c];
  }
  if (unique)
    for (dlong c = 0; c < Nids; ++c) {
      const hlong id = ids[local[c]];
      const hlong a = id < 0 ? -id : id;
      ids[local[c]] = isOwner[c] ? a : -a;
    }
  setupRowBlocks();

  o_rowStartsT = platform.malloc(rowStartsT);
  o_colIdsT = platform.malloc(colIdsT);
  o_rowStartsN = platform.malloc(rowStartsN);
  o_colIdsN = platform.malloc(colIdsN);
}

void ogs_t::setupRowBlocks() {
  const int perBlock = 512;  // include/ogs/ogsOperator.hpp:122-123
  const dlong NrowsT = NlocalT, NrowsN = NlocalP;
  std::vector<dlong> bT, bN;
  dlong sumT = 0, sumN = 0;
  if (NrowsT) bT.push_back(0);
  if (NrowsN) bN.push_back(0);
  for (dlong i = 0; i < NrowsT; ++i) {
    const dlong szN = rowStartsN[i + 1] - rowStartsN[i];
    const dlong szT = rowStartsT[i + 1] - rowStartsT[i];
    LIBP_ABORT("Multiplicity of global node id: " << i << " is too large", szN > perBlock || szT > perBlock);
    if (sumN + szN > perBlock) {
      bN.push_back(i);
      sumN = szN;
    } else {
      sumN += szN;
    }
    if (sumT + szT > perBlock) {
      bT.push_back(i);
      sumT = szT;
    } else {
      sumT += szT;
    }
  }
  // the reference's arrays hold block starts in [1..NrowBlocks) with entry 0 = 0
  NrowBlocksT = (dlong)bT.size();
  NrowBlocksN = (dlong)bN.size();
  blockRowStartsT.malloc(NrowBlocksT + 1, 0);
  blockRowStartsN.malloc(NrowBlocksN + 1, 0);
  for (dlong b = 1; b < NrowBlocksT; ++b) blockRowStartsT[b] = bT[b];
  for (dlong b = 1; b < NrowBlocksN; ++b) blockRowStartsN[b] = bN[b];
  blockRowStartsT[NrowBlocksT] = NrowsT;
  blockRowStartsN[NrowBlocksN] = NrowsN;
}

template <>
void ogs_t::Gather<double>(deviceMemory<double> o_gv, deviceMemory<double> o_v, const int, const Op,
                           const Transpose trans) {
  if (trans == Trans)
    LIBP_CALL(libp_ogs_gather_f64(NlocalT, o_rowStartsT.ptr(), o_colIdsT.ptr(), o_v.ptr(), o_gv.ptr(), platform.s()));
  else
    LIBP_CALL(libp_ogs_gather_f64(NlocalP, o_rowStartsN.ptr(), o_colIdsN.ptr(), o_v.ptr(), o_gv.ptr(), platform.s()));
}

template <>
void ogs_t::Gather<float>(deviceMemory<float> o_gv, deviceMemory<float> o_v, const int, const Op,
                          const Transpose trans) {
  if (trans == Trans)
    LIBP_CALL(libp_ogs_gather_f32(NlocalT, o_rowStartsT.ptr(), o_colIdsT.ptr(), o_v.ptr(), o_gv.ptr(), platform.s()));
  else
    LIBP_CALL(libp_ogs_gather_f32(NlocalP, o_rowStartsN.ptr(), o_colIdsN.ptr(), o_v.ptr(), o_gv.ptr(), platform.s()));
}

void ogs_t::Scatter(deviceMemory<dfloat> o_v, deviceMemory<dfloat> o_gv, const int, const Transpose trans) {
  if (trans == NoTrans)
    LIBP_CALL(libp_ogs_scatter_f64(NlocalT, o_rowStartsT.ptr(), o_colIdsT.ptr(), o_gv.ptr(), o_v.ptr(), platform.s()));
  else
    LIBP_CALL(libp_ogs_scatter_f64(NlocalP, o_rowStartsN.ptr(), o_colIdsN.ptr(), o_gv.ptr(), o_v.ptr(), platform.s()));
}

void ogs_t::SetupGlobalToLocalMapping(memory<dlong>& GlobalToLocal) const {
  GlobalToLocal.malloc(N, -1);
  for (dlong r = 0; r < NlocalT; ++r)
    for (dlong g = rowStartsT[r]; g < rowStartsT[r + 1]; ++g) GlobalToLocal[colIdsT[g]] = r;
}

}  // namespace ogs
}  // namespace libp
