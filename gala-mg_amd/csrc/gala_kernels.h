// Launch wrappers for the gfx950 kernels (host-callable, stream-ordered).
#pragma once
#include "gala_internal.h"
#include <algorithm>

namespace gala {

// ---- common (kernels_common.hip) ----
int64_t scan_tmp_elems(int64_t n);
template <class In, class V>
void scan_exclusive(const In* in, V* out, int64_t n, V* tmp, V* total, hipStream_t s);

int64_t part_tmp_elems(int64_t n);
void partition_bins(int mode, const int32_t* order, const int8_t* binid, const uint8_t* active,
                    int64_t n, int32_t* out, int32_t* bstart, int32_t* tmp, hipStream_t s);

template <class V>
void fill(V* p, int64_t n, V x, hipStream_t s);
void copy_i32(const int32_t* a, int32_t* b, int64_t n, hipStream_t s);
void reset_counters(Counters* c, long long min_init, double min_init_f, hipStream_t s);

// ---- Louvain phase kernels (kernels_louvain.hip) ----
// Graph view on the device.  off has n+1 exclusive offsets.
template <class T>
struct DevGraph {
  int32_t n;
  int64_t m;  // directed entries
  const int64_t* off;
  const int32_t* nbr;
  const typename T::w_t* w;  // nullptr => every weight is 1
};

template <class T>
struct DevState {
  typename T::tot_t *K, *Self, *Tot, *In, *dTot;
  int32_t *size, *dsize, *prev, *cur;
  uint8_t* active;
  int8_t* binid;
  int64_t* hub_off;  // per-vertex offset into the global hash arena (bin 8), in slots
  int32_t* hub_keys;
  typename T::acc_t* hub_vals;
  Counters* ctr;
};

template <class T>
void launch_init(const DevGraph<T>& g, DevState<T>& st, hipStream_t s);

// decide-and-move for one bin over front[bstart[b] .. bstart[b+1]) (device counts);
// `cap` = host upper bound of that range (used only to size the grid).
template <class T>
void launch_decide(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* front,
                   const int32_t* bstart, int cap, typename T::tot_t m2, hipStream_t s);

// Tot += dTot, size += dsize, sums of Tot^2, min Tot, non-empty count.
template <class T>
void launch_apply(DevState<T>& st, int32_t n, typename T::tot_t m2, hipStream_t s);

// In[v] and the prune flag over order[bstart[b] .. bstart[b+1]).
template <class T>
void launch_weight(int bin, const DevGraph<T>& g, DevState<T>& st, const int32_t* order,
                   const int32_t* bstart, int cap, typename T::tot_t m2, hipStream_t s);

// ---- contraction (kernels_contract.hip) ----
template <class T>
struct ContractBufs {
  int32_t* cnt;       // [n]   community sizes
  int32_t* nid;       // [n]   new id of community (exclusive scan of cnt > 0)
  int64_t* ecnt;      // [n+1] entries per new community
  int64_t* eoff;      // [n+1]
  int64_t* ecur;      // [n]
  int32_t* skey;      // [m]   staging keys
  typename T::acc_t* sval;  // [m]   staging values
  int32_t* dcnt;      // [n+1] distinct neighbours per new community
  int64_t* noff;      // [n+1]
  int32_t* big_list;  // [n]   communities taking the global-hash path
  int32_t* nbig;      // [1]
  int64_t* big_tab;   // [n+1] arena offsets (slots) of each large community's table
  int32_t* arena_keys;
  typename T::acc_t* arena_vals;
  uint32_t* arena_bits;
  int64_t arena_slots;
  int64_t arena_words;
  int32_t* i32_tmp;
  int64_t* i64_tmp;
  int32_t* i32_scalar;  // [4]
  int64_t* i64_scalar;  // [4]
};

template <class T>
struct ContractOut {
  int64_t* off;  // [C+1]
  int32_t* nbr;
  typename T::w_t* w;
};

// Renumber + member lists + per-community entry counts; returns C and total
// staged entries through host pointers (one sync).
template <class T>
int contract_prepare(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, int32_t* C_out,
                     int64_t* staged_out, hipStream_t s);
// Merge rows, canonical order, write the new CSR into out (sized by caller
// from the returned entry count).  Returns the new directed-entry count.
template <class T>
int contract_build(const DevGraph<T>& g, const int32_t* comm, ContractBufs<T>& cb, int32_t C,
                   int64_t staged, ContractOut<T>& out, int64_t* m_out, hipStream_t s);

// final[v] = nid[comm[final[v]]]
void compose_map(int32_t* fin, int32_t n0, const int32_t* comm, const int32_t* nid, hipStream_t s);

}  // namespace gala
