// Host-side construction of the element patches of the batched (many right-hand sides) element
// pass. Pure CPU code: callable without a GPU.
#include <cstdint>
#include <vector>
#include "../../include/exfem.h"

extern "C" {

// Patches of one element block, grown breadth-first through shared nodes: a patch starts at the
// lowest unassigned element and takes queued neighbours (elements sharing a node, in discovery
// order) while it has fewer than max_elems elements and its node count stays within nmax; an
// element that would overflow the cap is left for a later patch. Every element lands in exactly
// one patch. Outputs (caller-allocated): cptr [ne+1] element ranges of the patches, elem [ne]
// element ids patch after patch, lmap [ne*nnode] the patch-local row (< nmax) of every local node
// in that order, rbase [ne+1] first partial row of each patch (starting at row0) and rownode
// [ne*nnode] the node of every partial row (first rbase[nchunk]-row0 entries used).
// Returns the number of patches, or a negative status.
int64_t exfem_patch_build(int64_t ne, int nnode, int64_t nn, const int32_t* conn, int nmax,
                          int max_elems, int64_t row0, int32_t* cptr, int32_t* elem,
                          uint8_t* lmap, int32_t* rbase, int32_t* rownode) {
  if (ne < 0 || nn < 0 || nnode <= 0 || nmax < nnode || nmax > 255 || max_elems <= 0)
    return EXFEM_ERR_ARG;
  if (ne && (!conn || !cptr || !elem || !lmap || !rbase || !rownode)) return EXFEM_ERR_ARG;
  for (int64_t i = 0; i < ne * nnode; ++i)
    if (conn[i] < 0 || conn[i] >= nn) return EXFEM_ERR_ARG;
  std::vector<int64_t> eptr(nn + 1, 0);
  for (int64_t i = 0; i < ne * nnode; ++i) eptr[conn[i] + 1]++;
  for (int64_t n = 0; n < nn; ++n) eptr[n + 1] += eptr[n];
  std::vector<int32_t> eidx(eptr[nn] > 0 ? eptr[nn] : 1);
  {
    std::vector<int64_t> fill(eptr.begin(), eptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < nnode; ++a) eidx[fill[conn[e * nnode + a]]++] = (int32_t)e;
  }
  std::vector<char> assigned(ne, 0);
  std::vector<int64_t> nstamp(nn, -1), qstamp(ne, -1);
  std::vector<int32_t> lrow(nn, 0), queue;
  int64_t nchunk = 0, placed = 0, rows = 0;
  cptr[0] = 0;
  rbase[0] = (int32_t)row0;
  for (int64_t seed = 0; seed < ne; ++seed) {
    if (assigned[seed]) continue;
    const int64_t c = nchunk;
    queue.clear();
    queue.push_back((int32_t)seed);
    qstamp[seed] = c;
    size_t head = 0;
    int nodes = 0, members = 0;
    while (head < queue.size() && members < max_elems) {
      const int32_t e = queue[head++];
      if (assigned[e]) continue;
      int add = 0;
      for (int a = 0; a < nnode; ++a) add += nstamp[conn[(int64_t)e * nnode + a]] != c;
      if (nodes + add > nmax) continue;
      assigned[e] = 1;
      ++members;
      for (int a = 0; a < nnode; ++a) {
        const int32_t n = conn[(int64_t)e * nnode + a];
        if (nstamp[n] != c) {
          nstamp[n] = c;
          lrow[n] = nodes++;
          rownode[rows++] = n;
        }
        lmap[placed * nnode + a] = (uint8_t)lrow[n];
      }
      elem[placed++] = e;
      for (int a = 0; a < nnode; ++a) {
        const int32_t n = conn[(int64_t)e * nnode + a];
        for (int64_t k = eptr[n]; k < eptr[n + 1]; ++k) {
          const int32_t f = eidx[k];
          if (!assigned[f] && qstamp[f] != c) {
            qstamp[f] = c;
            queue.push_back(f);
          }
        }
      }
    }
    ++nchunk;
    cptr[nchunk] = (int32_t)placed;
    if (row0 + rows >= INT32_MAX) return EXFEM_ERR_ARG;
    rbase[nchunk] = (int32_t)(row0 + rows);
  }
  return nchunk;
}

}  // extern "C"
