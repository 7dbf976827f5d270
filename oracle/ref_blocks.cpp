// Test infrastructure, built only from the reference tree (make -C oracle ref):
// drives the reference's block-bitset layouts and their intersections
// (src/set/layouts/bitset_new.hpp:66-108, src/set/layouts/new_type.hpp:50-88,
// src/set/ops/intersection.hpp:1338-1649) over set pairs read from a file and
// writes every image and result to another file, so the oracle's restatement can
// be compared with the reference byte for byte.
//
// input:  repeated { u32 la; u32 a[la]; u32 lb; u32 b[lb] } (sorted, distinct)
// output: per pair, blobs { u64 card; u64 type; u64 nbytes; bytes[nbytes] } for
//         bitset_new(a), bitset_new(b), new_type(a), new_type(b),
//         bitset_new(a) & bitset_new(b), uint(a) & bitset_new(b), new_type(a) & new_type(b)
#include <cassert>

#define WRITE_VECTOR 1
#include "set/ops.hpp"

#include <cstdio>
#include <vector>

static void blob(FILE *out, uint64_t card, uint64_t type, const uint8_t *p, uint64_t n) {
  fwrite(&card, 8, 1, out);
  fwrite(&type, 8, 1, out);
  fwrite(&n, 8, 1, out);
  if (n) fwrite(p, 1, n, out);
}

static bool read_set(FILE *in, std::vector<uint32_t> &v) {
  uint32_t n = 0;
  if (fread(&n, 4, 1, in) != 1) return false;
  v.assign(n, 0);
  return n == 0 || fread(v.data(), 4, n, in) == n;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
  if (!in || !out) return 2;
  common::alloc_scratch_space(1 << 24, 1);
  std::vector<uint32_t> a, b;
  while (read_set(in, a) && read_set(in, b)) {
    const size_t la = a.size(), lb = b.size(), cap = 32 * (la + lb) + 256;
    std::vector<uint8_t> bna(cap), bnb(cap), nta(cap), ntb(cap), c1(cap), c2(cap), c3(cap);
    const size_t n_bna = bitset_new::build(bna.data(), a.data(), la);
    const size_t n_bnb = bitset_new::build(bnb.data(), b.data(), lb);
    const size_t n_nta = new_type::build(nta.data(), a.data(), la);
    const size_t n_ntb = new_type::build(ntb.data(), b.data(), lb);
    blob(out, la, common::BITSET_NEW, bna.data(), n_bna);
    blob(out, lb, common::BITSET_NEW, bnb.data(), n_bnb);
    blob(out, la, common::NEW_TYPE, nta.data(), n_nta);
    blob(out, lb, common::NEW_TYPE, ntb.data(), n_ntb);

    const Set<bitset_new> A(bna.data(), la, n_bna, common::BITSET_NEW), B(bnb.data(), lb, n_bnb, common::BITSET_NEW);
    Set<bitset_new> C1(c1.data());
    ops::set_intersect(&C1, &A, &B);
    blob(out, C1.cardinality, C1.type, C1.data, C1.number_of_bytes);

    const Set<uinteger> AU((uint8_t *)a.data(), la, 4 * la, common::UINTEGER);
    Set<uinteger> C2(c2.data());
    ops::set_intersect(&C2, &AU, &B);
    blob(out, C2.cardinality, C2.type, C2.data, C2.number_of_bytes);

    const Set<new_type> NA(nta.data(), la, n_nta, common::NEW_TYPE), NB(ntb.data(), lb, n_ntb, common::NEW_TYPE);
    Set<new_type> C3(c3.data());
    ops::set_intersect(&C3, &NA, &NB);
    blob(out, C3.cardinality, C3.type, C3.data, C3.number_of_bytes);
  }
  fclose(out);
  return 0;
}
