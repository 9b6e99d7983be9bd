// The workgroup order of the list passes (xcd_block in pbf_device.h), run
// on the host: for every grid of 1 .. 20,000 workgroups and for a few
// larger ones (up to the 4,096,000-particle grid of 16,000 and 2^20 + 7), the logical blocks are a permutation of [0, nb), and each
// XCD (hardware block mod 8) receives its logical blocks in runs of
// consecutive indices no shorter than min(nb / 8, 256) but for the tail.
#include <cstdio>
#include <vector>

#include "pbf_device.h"

static bool check(uint32_t nb) {
  std::vector<uint8_t> seen(nb, 0);
  for (uint32_t b = 0; b < nb; b++) {
    const uint32_t l = pbf::xcd_block(b, nb);
    if (l >= nb || seen[l]) {
      std::printf("nb %u: block %u -> %u (%s)\n", nb, b, l, l >= nb ? "out of range" : "twice");
      return false;
    }
    seen[l] = 1;
  }
  // Runs per XCD: successive hardware blocks of one XCD (b, b + 8, ...)
  // take consecutive logical blocks, except where a run ends.
  const uint32_t q = nb >> 3;
  const uint32_t run = q > pbf::kXcdRun ? pbf::kXcdRun : q;
  if (run > 1) {
    const uint32_t full = q > pbf::kXcdRun ? (nb / (8u * pbf::kXcdRun)) * (8u * pbf::kXcdRun) : 8u * q;
    for (uint32_t x = 0; x < 8; x++) {
      uint32_t len = 1;
      for (uint32_t b = x + 8; b < full; b += 8) {
        if (pbf::xcd_block(b, nb) == pbf::xcd_block(b - 8, nb) + 1) {
          len++;
        } else {
          if (len < run) {
            std::printf("nb %u: XCD %u has a run of %u < %u before block %u\n", nb, x, len, run, b);
            return false;
          }
          len = 1;
        }
      }
    }
  }
  return true;
}

int main() {
  for (uint32_t nb = 1; nb <= 20000; nb++)
    if (!check(nb)) return 1;
  for (uint32_t nb : {65535u, 65536u, 70001u, 128003u, 1u << 20, (1u << 20) + 7u})
    if (!check(nb)) return 1;
  std::printf("xcd order: all grids checked\n");
  return 0;
}
