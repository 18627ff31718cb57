"""Writes tests/golden/philox_rocrand.json: Philox4x32-10 vectors produced by
rocRAND's philox4x32_10 engine (the installed header
/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, compiled here for the host).

This is an implementation independent of the oracle's, used to pin
oracle.orc_philox4x32_10 and, through the GPU tests, the HIP noise kernel.
Run once in a container with ROCm; the JSON is committed.
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};
int main() {
    Probe p;
    unsigned s = 12345u;
    auto nxt = [&]() { s = s * 1664525u + 1013904223u; return s; };
    const unsigned fixed[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {0u, 0u, 0u, 0u, 42u, 0u},
        {1u, 0u, 0u, 0u, 42u, 0u},
        {65535u, 7u, 2u, 0u, 42u, 0u},
        {1048575u, 63u, 2u, 0u, 105u, 0u},
    };
    printf("[\n");
    int n = 0;
    for (auto& f : fixed) {
        uint4 r = p.run(make_uint4(f[0], f[1], f[2], f[3]), make_uint2(f[4], f[5]));
        printf("%s[[%u,%u,%u,%u],[%u,%u],[%u,%u,%u,%u]]\n", n++ ? "," : "", f[0], f[1], f[2], f[3],
               f[4], f[5], r.x, r.y, r.z, r.w);
    }
    for (int i = 0; i < 64; ++i) {
        unsigned c[4] = {nxt(), nxt(), nxt(), nxt()}, k[2] = {nxt(), nxt()};
        uint4 r = p.run(make_uint4(c[0], c[1], c[2], c[3]), make_uint2(k[0], k[1]));
        printf(",[[%u,%u,%u,%u],[%u,%u],[%u,%u,%u,%u]]\n", c[0], c[1], c[2], c[3], k[0], k[1],
               r.x, r.y, r.z, r.w);
    }
    printf("]\n");
    return 0;
}
"""


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.cpp")
        exe = os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.check_call(["hipcc", "-x", "c++", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                               src, "-o", exe, "-L/opt/rocm/lib", "-lamdhip64"])
        out = subprocess.check_output([exe]).decode()
    vecs = json.loads(out)
    doc = {
        "source": "rocRAND philox4x32_10_engine::ten_rounds (ROCm 7.2 header), host build",
        "format": "[counter[4], key[2], output[4]] as unsigned 32-bit integers",
        "vectors": vecs,
    }
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        f.write('{"source": %s,\n "format": %s,\n "vectors": [\n' % (json.dumps(doc["source"]),
                                                                   json.dumps(doc["format"])))
        f.write(",\n".join("  " + json.dumps(v) for v in vecs))
        f.write("\n]}\n")
    print("wrote", len(vecs), "vectors")


if __name__ == "__main__":
    main()
