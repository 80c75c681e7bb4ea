// Writes Philox4x32-10 vectors from rocRAND's independent implementation
// (rocrand_philox4x32_10.h, host path) for the counters/keys given on stdin:
// each line "c0 c1 c2 c3 k0 k1" (hex) -> "c0 c1 c2 c3 k0 k1 o0 o1 o2 o3".
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    unsigned c[4], k[2];
    Probe p;
    while (std::scanf("%x %x %x %x %x %x", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
        uint4 o = p.run(make_uint4(c[0], c[1], c[2], c[3]), make_uint2(k[0], k[1]));
        std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c[0], c[1], c[2], c[3], k[0], k[1],
                    o.x, o.y, o.z, o.w);
    }
    return 0;
}
