"""Generate hitx_normal_tables.h: the 256-level ziggurat tables of NumPy's
Generator.standard_normal (the forcing noise stream of the reference,
OOP/forcing.py:66,77 with numpy.random.default_rng).

The tables follow the published construction (Marsaglia & Tsang, "The
Ziggurat Method for Generating Random Variables", J. Stat. Softw. 5(8),
2000) with 256 layers, 52-bit mantissa draws, tail start r and layer area v
as NumPy uses them.  The construction is done in extended precision; the
layer widths wi[] are then pinned black-box to the doubles that reproduce
Generator.standard_normal bit for bit: for each layer, the double w with
fl(rabs * w) == |x| for every fast-path draw (rabs, x) observed from the raw
PCG64 words (bit_generator.random_raw) of the same seeds.  ki[] and fi[] only
take part in comparisons and stay as constructed (a draw lands within their
rounding of a boundary with probability ~1e-14).

Run from this directory:  python gen_normal_tables.py
"""

from __future__ import annotations

import numpy as np

R = "3.6541528853610088"
V = "4.92867323397465524494e-3"
SEEDS = (1235, 11, 77, 5, 6, 2024)
N_DRAWS = 400_000


def construct():
    LD = np.longdouble
    dn = LD(R)
    tn = dn
    vn = LD(V)
    m1 = LD(2) ** 52
    ki = [0] * 256
    wi = [0.0] * 256
    fi = [0.0] * 256
    q = vn / np.exp(LD(-0.5) * dn * dn)
    ki[0] = int((dn / q) * m1)
    ki[1] = 0
    wi[0] = float(q / m1)
    wi[255] = float(dn / m1)
    fi[0] = 1.0
    fi[255] = float(np.exp(LD(-0.5) * dn * dn))
    for i in range(254, 0, -1):
        dn = np.sqrt(LD(-2) * np.log(vn / dn + np.exp(LD(-0.5) * dn * dn)))
        ki[i + 1] = int((dn / tn) * m1)
        tn = dn
        fi[i] = float(np.exp(LD(-0.5) * dn * dn))
        wi[i] = float(dn / m1)
    return ki, wi, fi


def next_double(w):
    return float(int(w) >> 11) * (1.0 / 9007199254740992.0)


def collect(ki, wi, fi):
    """(layer, rabs, |x|) of first-attempt outputs, parsed from the raw words."""
    rr = float(R)
    samples = {i: ([], []) for i in range(256)}
    for seed in SEEDS:
        words = np.random.default_rng(seed).bit_generator.random_raw(int(N_DRAWS * 1.1) + 64)
        xs = np.random.default_rng(seed).standard_normal(N_DRAWS)
        idx_all = (words & np.uint64(0xFF)).astype(np.int64)
        rabs_all = ((words >> np.uint64(9)) & np.uint64((1 << 52) - 1))
        p = 0
        for k in range(N_DRAWS):
            while True:
                idx = int(idx_all[p])
                rabs = int(rabs_all[p])
                x = rabs * wi[idx]
                if rabs < ki[idx]:
                    samples[idx][0].append(rabs)
                    samples[idx][1].append(abs(xs[k]))
                    p += 1
                    break
                if idx == 0:
                    q = p + 1
                    while True:
                        xx = -(1.0 / rr) * np.log1p(-next_double(words[q]))
                        yy = -np.log1p(-next_double(words[q + 1]))
                        q += 2
                        if yy + yy > xx * xx:
                            break
                    p = q
                    break
                u = next_double(words[p + 1])
                p += 2
                if (fi[idx - 1] - fi[idx]) * u + fi[idx] < np.exp(-0.5 * x * x):
                    if abs(abs(xs[k]) - x) <= 1e-12 * abs(x):
                        samples[idx][0].append(rabs)
                        samples[idx][1].append(abs(xs[k]))
                    break
    return samples


def pin(wi, samples):
    out = list(wi)
    for i in range(256):
        ra = np.array(samples[i][0], dtype=np.float64)
        xa = np.array(samples[i][1], dtype=np.float64)
        if ra.size < 16:
            raise SystemExit(f"too few samples for layer {i}")
        best = None
        for direction in (-np.inf, np.inf):
            c = np.float64(wi[i])
            for _ in range(400):
                if np.all(ra * c == xa):
                    best = c
                    break
                c = np.nextafter(c, direction)
            if best is not None:
                break
        if best is None:
            raise SystemExit(f"no double reproduces layer {i}")
        out[i] = float(best)
    return out


def main():
    ki, wi, fi = construct()
    wi = pin(wi, collect(ki, wi, fi))
    lines = [
        "// Generated by gen_normal_tables.py: 256-level ziggurat tables of the",
        "// standard normal (Marsaglia & Tsang 2000 construction; layer widths pinned",
        "// to NumPy Generator.standard_normal).  Do not edit.",
        "#pragma once",
        "#include <stdint.h>",
        "namespace hitx {",
        f"constexpr double kZigR = {float(R).hex()};  // {R}",
        f"constexpr double kZigInvR = {(1.0 / float(R)).hex()};",
        "__device__ const uint64_t kZigKi[256] = {",
    ]
    lines += ["    " + ", ".join(f"0x{v:013x}ull" for v in ki[j:j + 4]) + "," for j in range(0, 256, 4)]
    lines += ["};", "__device__ const double kZigWi[256] = {"]
    lines += ["    " + ", ".join(v.hex() for v in wi[j:j + 3]) + "," for j in range(0, 256, 3)]
    lines += ["};", "__device__ const double kZigFi[256] = {"]
    lines += ["    " + ", ".join(v.hex() for v in fi[j:j + 3]) + "," for j in range(0, 256, 3)]
    lines += ["};", "}  // namespace hitx", ""]
    with open("hitx_normal_tables.h", "w") as f:
        f.write("\n".join(lines))
    print("wrote hitx_normal_tables.h")


if __name__ == "__main__":
    main()
