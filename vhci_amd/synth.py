"""Seeded synthetic force fields in the reference's input format.

The generator is deterministic across machines: a splitmix64-seeded
xoshiro256** stream (Blackman & Vigna) in pure Python integer arithmetic, so
every box that runs it writes byte-identical input files.

Workloads (seed = 0x5648434900 + config):
  c0  conventional VCI: 12 modes, all cubic i<=j<=k (364) and all
      semi-diagonal quartic iiii/iiij/iijj (210); caps 4, Ntot 4 (N = 1820);
      eps1 = 0, Nstate = 50, no PT2.
  c1  heat-bath VCI + PT2: 15 modes, all 680 cubic, the 330 semi-diagonal
      quartics plus 270 distinct f_iijk (600 quartic); Ntot 2, caps 6,
      eps1 = 1.0, Nstate = 100, eps2 = 0.01.
  c2  48 modes, frequencies in [150, 3200]: all 19,600 cubic f_ijk with |f|
      log-uniform in [0.5, 60 w/3200], and the 3,432 semi-diagonal plus
      51,888 f_iijk quartics with |f| log-uniform in [0.25, 12 w/3200], w
      the geometric mean frequency of the constant's modes; |f| < 0.5 is
      dropped (42,212 quartics kept) and f_iiii > 0.  Ntot 2, caps 6,
      eps1 = 1, Nstate = 128, eps2 = 0.2.  The workload fixes the counts
      (about 2e4 cubic, 4e4 quartic), the 0.5 cutoff and log-uniform
      magnitudes, not the upper ends.  With config 0's ends (300 and 60 for
      every mode) the 48-mode heat bath does not converge (the basis passes
      16 M states): every mode with a few hundred cm^-1 is then coupled as
      strongly as the 3000 cm^-1 ones.  Ends proportional to the modes'
      frequencies keep each mode about as anharmonic relative to its own
      frequency, as in real force fields, and the basis converges after
      three iterations at 1,598,434 states (1225 -> 1,003,305 ->
      1,548,210 -> 1,598,434) with about 1.7e9 nonzeros.
  c4  sextic stress case, 60 modes, frequencies in [100, 3300]: cubic and
      quartic constants as for c2 (all 37,820 cubic f_ijk, and the
      semi-diagonal plus every other f_iijk quartic with |f| >= 0.5 kept:
      43,722), plus all 7,140 semi-diagonal quintic (i^5, i^4 j, i^3 j^2)
      and 8,910 semi-diagonal sextic (i^6, i^5 j, i^4 j^2, i^3 j^3)
      constants with |f| log-uniform in [0.01, 5] and f_i^6 > 0: 97,592
      constants of orders 3-6.  Ntot 2, caps 6, eps1 = 0.5, Nstate = 200,
      eps2 = 0.05.
  c3  serving stream: request r (seed 0x5648434900 + 1000 + r) draws
      M uniform in [8, 36], eps1 from {0.5, 1, 2}, Nstate uniform in
      [20, 200] and eps2 from {0, 0.01, 0.1}; the field has all cubic and
      all semi-diagonal quartic constants with frequencies in [700, 3300]
      as for c0, but cubic |f| in [1, 100] and f_iiii > 0: with c0's
      cubic range and random quartic signs, many small-M draws have a
      potential unbounded below, and heat bath (which caps no quanta)
      then grows the basis into the collapse without end.  Ntot 2, caps 6,
      PT2 on.  The ranges can be narrowed for short runs.
Frequencies are uniform in [lo, hi]; |f| is log-uniform in [a, b] with a
random sign.  Values are written with 9 decimals.
"""

MASK = (1 << 64) - 1
BASE_SEED = 0x5648434900


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK


class Rng:
    """xoshiro256** seeded by splitmix64."""

    def __init__(self, seed):
        st = seed & MASK
        self.s = []
        for _ in range(4):
            st, z = _splitmix(st)
            self.s.append(z)

    def next(self):
        s = self.s
        result = (_rotl((s[1] * 5) & MASK, 7) * 9) & MASK
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def logmag(self, a, b):
        import math
        mag = 10 ** self.uniform(math.log10(a), math.log10(b))
        sign = -1.0 if (self.next() >> 63) else 1.0
        return sign * mag

    def randint(self, n):
        return self.next() % n


def _cubic_all(M):
    return [(i, j, k) for i in range(M) for j in range(i, M) for k in range(j, M)]


def _semidiag_quartic(M):
    out = []
    for i in range(M):
        out.append((i, i, i, i))
    for i in range(M):
        for j in range(M):
            if i != j:
                out.append((i, i, i, j))
    for i in range(M):
        for j in range(i + 1, M):
            out.append((i, i, j, j))
    return out


def _iijk_pool(M):
    out = []
    for i in range(M):
        for j in range(M):
            for k in range(j + 1, M):
                if j != i and k != i:
                    out.append((i, i, j, k))
    return out


def _semidiag(M, order):
    """Constants of one order over at most two distinct modes, i^a j^b with
    a >= b: i^order, then i^a j^b for i != j (a > b) or i < j (a == b)."""
    out = []
    for a in range(order, (order - 1) // 2, -1):
        b = order - a
        if b == 0:
            out += [(i,) * order for i in range(M)]
        elif a > b:
            out += [(i,) * a + (j,) * b for i in range(M) for j in range(M) if i != j]
        else:
            out += [(i,) * a + (j,) * b for i in range(M) for j in range(i + 1, M)]
    return out


def write_input(path_or_none, comment, eps1, nstate, pt2, eps2, ntot, freqs, caps, fcs):
    lines = ["Comment: " + comment,
             "HCI_Eps: %s" % repr(float(eps1)),
             "Nstate: %d" % nstate,
             "PT2: %d" % (1 if pt2 else 0),
             "PT2_Eps: %s" % repr(float(eps2)),
             "Max_quanta: %d" % ntot,
             "Modes: %d" % len(freqs)]
    for i, (f, c) in enumerate(zip(freqs, caps)):
        lines.append(" %d %.6f %d" % (i, f, c))
    lines.append("Force_constants: %d" % len(fcs))
    for modes, v in fcs:
        lines.append(" %d %s %.9f" % (len(modes), " ".join(str(m) for m in modes), v))
    text = "\n".join(lines) + "\n"
    if path_or_none:
        with open(path_or_none, "w") as f:
            f.write(text)
    return text


def field(M, flo, fhi, seed, cubic=(1.0, 300.0), quartic=(0.1, 60.0), n_extra_iijk=0, cutoff=0.0,
          cubic_pool=None, quartic_pool=None, positive_iiii=False, freq_scale=0.0, higher=()):
    """Frequencies uniform in [flo, fhi]; one log-uniform magnitude with a
    random sign per constant of the pools.  freq_scale = p > 0 scales the
    upper end of each constant's range by (w / fhi)^p, w the geometric mean
    frequency of its modes (a range that would end below its lower end ends
    just above it).  Cubic and quartic constants below `cutoff` are dropped.
    higher: (pool, range) pairs of further constants (quintic, sextic) drawn
    after the quartics with no cutoff; their all-diagonal even constants are
    positive."""
    import math
    rng = Rng(seed)
    freqs = [rng.uniform(flo, fhi) for _ in range(M)]
    fcs = []

    def draw(t, rng_ab):
        a, b = rng_ab
        if freq_scale:
            w = math.exp(sum(math.log(freqs[i]) for i in t) / len(t))
            b = max(a * 1.0001, b * (w / fhi) ** freq_scale)
        return rng.logmag(a, b)

    for t in (cubic_pool if cubic_pool is not None else _cubic_all(M)):
        v = draw(t, cubic)
        if abs(v) >= cutoff:
            fcs.append((t, v))
    qp = quartic_pool if quartic_pool is not None else _semidiag_quartic(M)
    for t in qp:
        v = draw(t, quartic)
        if positive_iiii and t[0] == t[1] == t[2] == t[3]:
            v = abs(v)
        if abs(v) >= cutoff:
            fcs.append((t, v))
    for pool, rr in higher:
        for t in pool:
            v = draw(t, rr)
            if len(t) % 2 == 0 and len(set(t)) == 1:
                v = abs(v)
            fcs.append((t, v))
    if n_extra_iijk:
        pool = _iijk_pool(M)
        chosen = set()
        while len(chosen) < n_extra_iijk:
            chosen.add(rng.randint(len(pool)))
        for idx in sorted(chosen):
            fcs.append((pool[idx], rng.logmag(*quartic)))
    return freqs, fcs


def config_input(c, path=None, **override):
    """Input text (and file when path is given) of synthetic workload c."""
    seed = BASE_SEED + c
    if c == 0:
        M = 12
        freqs, fcs = field(M, 700.0, 3300.0, seed)
        args = dict(eps1=0.0, nstate=50, pt2=False, eps2=0.0, ntot=4, caps=[4] * M)
    elif c == 1:
        M = 15
        freqs, fcs = field(M, 700.0, 3300.0, seed, n_extra_iijk=270)
        args = dict(eps1=1.0, nstate=100, pt2=True, eps2=0.01, ntot=2, caps=[6] * M)
    elif c == 2:
        M = 48
        freqs, fcs = field(M, 150.0, 3200.0, seed, cubic=(0.5, 60.0), quartic=(0.25, 12.0), cutoff=0.5,
                           quartic_pool=_semidiag_quartic(M) + _iijk_pool(M), positive_iiii=True, freq_scale=1.0)
        args = dict(eps1=1.0, nstate=128, pt2=True, eps2=0.2, ntot=2, caps=[6] * M)
    elif c == 4:
        M = 60
        freqs, fcs = field(M, 100.0, 3300.0, seed, cubic=(0.5, 60.0), quartic=(0.25, 12.0), cutoff=0.5,
                           quartic_pool=_semidiag_quartic(M) + _iijk_pool(M)[::2], positive_iiii=True,
                           freq_scale=1.0, higher=((_semidiag(M, 5), (0.01, 5.0)), (_semidiag(M, 6), (0.01, 5.0))))
        args = dict(eps1=0.5, nstate=200, pt2=True, eps2=0.05, ntot=2, caps=[6] * M)
    else:
        raise ValueError("unknown workload %r" % c)
    args.update(override)
    caps = args.pop("caps")
    return write_input(path, "synthetic workload c%d seed %#x" % (c, seed), freqs=freqs, caps=caps, fcs=fcs,
                       **args)


def request_params(r, m_range=(8, 36), eps1_set=(0.5, 1.0, 2.0), nstate_range=(20, 200),
                   eps2_set=(0.0, 0.01, 0.1), eps2_zero_max_modes=None):
    """Parameters of serving request r: (M, eps1, Nstate, eps2, field seed).

    eps2 = 0 makes PT2 run over every connected state.  On fields of more
    than about 20 modes that space holds 10^8 to 10^9 states (1.4e8 for
    request 3, 25 modes, 43 s on one MI355X; request 11, 35 modes, ran for
    over 3 minutes), so a stream of seconds-long requests would be set by a
    handful of them.  With eps2_zero_max_modes = m, an eps2 = 0 draw on a
    field of more than m modes becomes eps2 = 0.01 (the same random draws
    otherwise, so every other request is unchanged)."""
    seed = BASE_SEED + 1000 + r
    rng = Rng(seed)
    M = m_range[0] + rng.randint(m_range[1] - m_range[0] + 1)
    eps1 = eps1_set[rng.randint(len(eps1_set))]
    nstate = nstate_range[0] + rng.randint(nstate_range[1] - nstate_range[0] + 1)
    eps2 = eps2_set[rng.randint(len(eps2_set))]
    if eps2 == 0.0 and eps2_zero_max_modes is not None and M > eps2_zero_max_modes:
        eps2 = 0.01
    return M, eps1, nstate, eps2, (seed + (1 << 40)) & MASK


def request_input(r, path=None, **ranges):
    """Input text of serving request r (workload c3)."""
    M, eps1, nstate, eps2, fseed = request_params(r, **ranges)
    freqs, fcs = field(M, 700.0, 3300.0, fseed, cubic=(1.0, 100.0), positive_iiii=True)
    return write_input(path, "synthetic serving request %d seed %#x" % (r, BASE_SEED + 1000 + r), eps1=eps1,
                       nstate=nstate, pt2=True, eps2=eps2, ntot=2, freqs=freqs, caps=[6] * M, fcs=fcs)


if __name__ == "__main__":
    import sys
    c = int(sys.argv[1])
    out = sys.argv[2]
    config_input(c, out)
