"""Deterministic workload definitions for the batching and serving runs.

splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) drives every random choice, so a workload is a pure
function of its seed.

batch_manifest():  96 maps, seeds 100..195.  For the map of seed s,
    x = splitmix64(s); side = (128, 192, 256, 384, 512, 768, 1024)[x % 7];
    iterations = 100 + (x >> 8) % 301; if (x >> 20) % 4 == 0 the map is
    rectangular, nx = side / 2, ny = side (aspect 1:2).
serve_requests(): Poisson arrivals (rate 200/s) from seed 12345: each request
    draws four uniforms u1..u4 in [0, 1): gap = -ln(1 - u1) / rate;
    side = 64 * round(exp(ln 256 + u2 * (ln 4096 - ln 256)) / 64);
    iterations = 50 + floor(u3 * 451); kind = vpipes+thermal if u4 < 0.7,
    else thermal.
"""
import math

MASK64 = (1 << 64) - 1
SIDES = (128, 192, 256, 384, 512, 768, 1024)


def splitmix64(x):
    """The splitmix64 output for state x (state advanced by the golden gamma)."""
    z = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        out = splitmix64(self.state)
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        return out

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def batch_manifest(n_maps=96, first_seed=100):
    maps = []
    for s in range(first_seed, first_seed + n_maps):
        x = splitmix64(s)
        side = SIDES[x % 7]
        iters = 100 + (x >> 8) % 301
        rect = (x >> 20) % 4 == 0
        maps.append({"seed": s, "nx": side // 2 if rect else side, "ny": side,
                     "iterations": int(iters)})
    return maps


def serve_requests(n=2000, rate=200.0, seed=12345, frac_vpipes=0.7):
    g = SplitMix64(seed)
    t = 0.0
    out = []
    lo, hi = math.log(256.0), math.log(4096.0)
    for i in range(n):
        u1, u2, u3, u4 = g.uniform(), g.uniform(), g.uniform(), g.uniform()
        t += -math.log(1.0 - u1) / rate
        side = 64 * int(round(math.exp(lo + u2 * (hi - lo)) / 64.0))
        side = min(max(side, 256), 4096)
        out.append({"id": i, "arrival_s": t, "side": side, "iterations": 50 + int(u3 * 451),
                    "kind": "vpipes+thermal" if u4 < frac_vpipes else "thermal",
                    "seed": 1000 + i})
    return out
