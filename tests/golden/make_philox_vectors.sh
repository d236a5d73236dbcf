#!/bin/sh
# Writes philox_rocrand_vectors.json: Philox4x32-10 outputs for pseudo-random
# (counter, key) pairs computed by the ROCm rocRAND header
# (rocrand_philox4x32_10.h, an implementation independent of this project),
# called on the host.  Used by tests/test_oracle.py to pin the oracle's Philox.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TMP=$(mktemp -d)
cat > "$TMP/gen.cpp" <<'CPP'
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>
struct Open : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    Open e;
    uint64_t s = 0x243F6A8885A308D3ull;  // splitmix64 stream of inputs
    auto next = [&]() { s += 0x9E3779B97F4A7C15ull; uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31); };
    std::printf("[\n");
    for (int i = 0; i < 64; ++i) {
        uint64_t a = next(), b = next(), k = next();
        uint4 c = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
        uint2 kk = {(unsigned)k, (unsigned)(k >> 32)};
        uint4 o = e.block(c, kk);
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c.x, c.y, c.z, c.w, kk.x, kk.y, o.x, o.y, o.z, o.w, i < 63 ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
CPP
/opt/rocm/bin/hipcc -O1 -o "$TMP/gen" "$TMP/gen.cpp"
"$TMP/gen" > "$HERE/philox_rocrand_vectors.json"
rm -rf "$TMP"
