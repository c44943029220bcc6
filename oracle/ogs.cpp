// CPU oracle (test infrastructure only): single-rank signed gather-scatter
// setup and operators, restated from libs/ogs/ogsSetup.cpp and ogsOperator.cpp.
#include <algorithm>
#include <cstdlib>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

namespace {
// include/ogs/ogsUtils.hpp:36-47
struct parallelNode_t {
  dlong localId;
  hlong baseId;
  dlong newId;
  int sign;
  int rank;
  int destRank;
};
inline hlong habs(hlong a) { return a < 0 ? -a : a; }
}  // namespace

/* ogsBase_t::Setup (ogsSetup.cpp:69-190) for kind=Signed on one rank, with
 * FindSharedNodes (:192-331), ConstructSharedNodes (:333-566) and
 * LocalSignedSetup (:569-683). The owner of every id group is drawn with
 * rand() % groupsize after an unstable std::sort by |baseId| (:241-275); this
 * restatement uses the same sort on the same key sequence and glibc's rand()
 * stream, which is the single-threaded path of the reference. */
void Ogs::Setup(dlong N_, std::vector<hlong>& ids, bool unique, GlibcRand& rng) {
  N = N_;
  dlong Nids = 0;
  for (dlong n = 0; n < N; n++)
    if (ids[n] != 0) Nids++;
  std::vector<parallelNode_t> nodes(Nids);
  Nids = 0;
  for (dlong n = 0; n < N; n++)
    if (ids[n] != 0) {
      nodes[Nids].localId = Nids;
      nodes[Nids].baseId = ids[n];
      nodes[Nids].rank = 0;
      nodes[Nids].destRank = 0;
      Nids++;
    }

  // FindSharedNodes: send order == local order on one rank.
  {
    std::vector<parallelNode_t> recv = nodes;
    for (dlong n = 0; n < Nids; ++n) recv[n].newId = n;
    std::sort(recv.begin(), recv.end(),
              [](const parallelNode_t& a, const parallelNode_t& b) { return habs(a.baseId) < habs(b.baseId); });
    dlong start = 0;
    for (dlong n = 0; n < Nids; n++) {
      if (n == Nids - 1 || habs(recv[n].baseId) != habs(recv[n + 1].baseId)) {
        dlong end = n + 1;
        if (unique) {
          const hlong baseId = habs(recv[start].baseId);
          const int m = (rng.next() % (end - start));
          for (dlong i = start; i < end; i++) recv[i].baseId = -baseId;
          recv[start + m].baseId = baseId;
        }
        for (dlong i = start; i < end; i++) recv[i].sign = 1;  // never shared across ranks
        start = n + 1;
      }
    }
    for (dlong n = 0; n < Nids; ++n) nodes[recv[n].newId] = recv[n];
  }

  // ConstructSharedNodes: groups in |baseId| order, numbered by first
  // encounter in local order (:409-433).
  std::vector<hlong> keys(Nids);
  for (dlong n = 0; n < Nids; ++n) keys[n] = habs(nodes[n].baseId);
  std::vector<hlong> uniq = keys;
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  const dlong NbaseIds = (dlong)uniq.size();
  std::vector<dlong> indexMap(NbaseIds, -1);
  NlocalT = NlocalP = 0;
  // every group has a leading positive id when unique==true (sign = +1)
  for (dlong n = 0; n < Nids; ++n) {
    dlong g = (dlong)(std::lower_bound(uniq.begin(), uniq.end(), keys[n]) - uniq.begin());
    if (indexMap[g] == -1) indexMap[g] = NlocalP++;
    nodes[n].newId = indexMap[g];
  }
  NlocalT = NlocalP;
  Ngather = NlocalP;

  // restore real local ids and write the owner signs back (:137-147)
  Nids = 0;
  for (dlong n = 0; n < N; n++)
    if (ids[n] != 0) {
      nodes[Nids].localId = n;
      if (unique) ids[n] = nodes[Nids].baseId;
      Nids++;
    }

  // LocalSignedSetup (:569-683)
  OgsOperator& G = gatherLocal;
  G.Ncols = N;
  G.NrowsN = NlocalP;
  G.NrowsT = NlocalT;
  std::vector<dlong> cntN(NlocalT, 0), cntT(NlocalT, 0);
  for (dlong i = 0; i < Nids; i++) {
    const dlong gid = nodes[i].newId;
    if (nodes[i].baseId > 0) cntN[gid]++;
    cntT[gid]++;
  }
  G.rowStartsN.assign(NlocalT + 1, 0);
  G.rowStartsT.assign(NlocalT + 1, 0);
  for (dlong i = 0; i < NlocalT; i++) {
    G.rowStartsN[i + 1] = G.rowStartsN[i] + cntN[i];
    G.rowStartsT[i + 1] = G.rowStartsT[i] + cntT[i];
    cntN[i] = 0;
    cntT[i] = 0;
  }
  G.colIdsN.assign(G.rowStartsN[NlocalT], 0);
  G.colIdsT.assign(G.rowStartsT[NlocalT], 0);
  for (dlong i = 0; i < Nids; i++) {
    const dlong gid = nodes[i].newId;
    if (nodes[i].baseId > 0) G.colIdsN[G.rowStartsN[gid] + cntN[gid]++] = nodes[i].localId;
    G.colIdsT[G.rowStartsT[gid] + cntT[gid]++] = nodes[i].localId;
  }
  G.setupRowBlocks();
}

/* ogsOperator.cpp:495-560. */
void OgsOperator::setupRowBlocks() {
  for (int pass = 0; pass < 2; ++pass) {
    dlong blockSumN = 0, blockSumT = 0;
    NrowBlocksN = 0;
    NrowBlocksT = 0;
    if (NrowsN) NrowBlocksN++;
    if (NrowsT) NrowBlocksT++;
    for (dlong i = 0; i < NrowsT; i++) {
      const dlong rowSizeN = rowStartsN[i + 1] - rowStartsN[i];
      const dlong rowSizeT = rowStartsT[i + 1] - rowStartsT[i];
      if (rowSizeN > gatherNodesPerBlock || rowSizeT > gatherNodesPerBlock)
        throw std::runtime_error("Multiplicity of global node id too large");
      if (blockSumN + rowSizeN > gatherNodesPerBlock) {
        if (pass) blockRowStartsN[NrowBlocksN] = i;
        NrowBlocksN++;
        blockSumN = rowSizeN;
      } else {
        blockSumN += rowSizeN;
      }
      if (blockSumT + rowSizeT > gatherNodesPerBlock) {
        if (pass) blockRowStartsT[NrowBlocksT] = i;
        NrowBlocksT++;
        blockSumT = rowSizeT;
      } else {
        blockSumT += rowSizeT;
      }
    }
    if (!pass) {
      blockRowStartsN.assign(NrowBlocksN + 1, 0);
      blockRowStartsT.assign(NrowBlocksT + 1, 0);
    }
  }
  blockRowStartsN[NrowBlocksN] = NrowsN;
  blockRowStartsT[NrowBlocksT] = NrowsT;
}

/* Host gather with Op=Add (ogsOperator.cpp:64-113): sum from zero in
 * ascending column order. trans=true uses the T (all copies) rows. */
template <typename T>
void OgsOperator::GatherAdd(T* gv, const T* v, bool trans, bool parallel) const {
  const dlong Nrows = trans ? NrowsT : NrowsN;
  const dlong* rs = trans ? rowStartsT.data() : rowStartsN.data();
  const dlong* ci = trans ? colIdsT.data() : colIdsN.data();
  // rows are independent: the reference's host loop runs them under OpenMP
  // (ogsOperator.cpp:87); each row still sums in ascending column order
#pragma omp parallel for schedule(static) if (parallel)
  for (dlong n = 0; n < Nrows; ++n) {
    T val = 0;
    for (dlong g = rs[n]; g < rs[n + 1]; ++g) val += v[ci[g]];
    gv[n] = val;
  }
}

/* Host scatter (ogsOperator.cpp:225-266): NoTrans writes every copy. */
template <typename T>
void OgsOperator::Scatter(T* v, const T* gv, bool trans) const {
  const dlong Nrows = trans ? NrowsN : NrowsT;
  const dlong* rs = trans ? rowStartsN.data() : rowStartsT.data();
  const dlong* ci = trans ? colIdsN.data() : colIdsT.data();
  for (dlong n = 0; n < Nrows; ++n)
    for (dlong g = rs[n]; g < rs[n + 1]; ++g) v[ci[g]] = gv[n];
}

template void OgsOperator::GatherAdd<double>(double*, const double*, bool, bool) const;
template void OgsOperator::GatherAdd<float>(float*, const float*, bool, bool) const;
template void OgsOperator::Scatter<double>(double*, const double*, bool) const;
template void OgsOperator::Scatter<float>(float*, const float*, bool) const;
template void OgsOperator::Scatter<dlong>(dlong*, const dlong*, bool) const;

/* ogs_t::SetupGlobalToLocalMapping (ogsSetup.cpp:867-886). */
void Ogs::SetupGlobalToLocalMapping(std::vector<dlong>& G2L) const {
  G2L.assign(N, -1);
  std::vector<dlong> idsv(NlocalT);
  for (dlong n = 0; n < NlocalT; ++n) idsv[n] = n;
  gatherLocal.Scatter<dlong>(G2L.data(), idsv.data(), false);
}

}  // namespace oracle
