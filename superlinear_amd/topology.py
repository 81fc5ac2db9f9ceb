"""SUS2-SH model topology generator (the ``init-sh`` step) and model writer.

Builds the spherical-harmonic product graph of a Static or coordination-Effect
model and writes it in the SUS2-SH text format, so that every benchmark and
test model can be generated from a seed.  The graph conventions follow the
reference generator (dev_src/sh_model_init.cpp):

* channel order (l from L down to 0, then k from K-1 down to 0) and one node
  per (l, m) for each channel (:671-691 Build, :582 BasicTensor);
* real-basis Clebsch-Gordan couplings (:439-536), products keyed and merged
  by (left, right, target) with left <= right (:599-657, :970-985);
* scalar enumeration per body order 2..6 with ``q-total`` (cross) or
  ``same-k`` (diagonal) factor pruning (:702-838); a same-k rank graph is a
  single-channel plan replicated per channel (:936-990, :1500-1540);
* segment layout raw basics | gate rank graph | outer rank graph, and the
  ``*_begin``/``*_product_*`` fields that mark them (:1490-1800).

Coefficient initialisation is not the reference's (its random stream is
``std::mt19937_64``, dev_src/mtpr.cpp:4082); values are drawn from numpy with
a caller seed.
"""
from math import factorial, sqrt

import numpy as np

EPS = 1.0e-12


def _triangle(a, b, c):
    return abs(a - b) <= c <= a + b


def clebsch_gordan(j1, m1, j2, m2, j, m):
    """Complex-basis Clebsch-Gordan <j1 m1 j2 m2 | j m> (Racah's formula)."""
    if m != m1 + m2 or not _triangle(j1, j2, j) or abs(m1) > j1 or abs(m2) > j2 or abs(m) > j:
        return 0.0
    pre = sqrt((2 * j + 1) * factorial(j1 + j2 - j) * factorial(j1 - j2 + j) * factorial(-j1 + j2 + j)
               / factorial(j1 + j2 + j + 1))
    pre *= sqrt(factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) * factorial(j2 - m2)
                * factorial(j + m) * factorial(j - m))
    s = 0.0
    for k in range(0, j1 + j2 + j + 1):
        den = [k, j1 + j2 - j - k, j1 - m1 - k, j2 + m2 - k, j - j2 + m1 + k, j - j1 - m2 + k]
        if min(den) < 0:
            continue
        p = 1.0
        for d in den:
            p *= factorial(d)
        s += (-1) ** k / p
    return pre * s


def _real_from_complex(rm, cm):
    """<complex Y_l^cm | real Y_l,rm>: the unitary map from the Condon-Shortley
    complex harmonics to the real harmonics of the EFS kernels."""
    if rm == 0:
        return 1.0 + 0j if cm == 0 else 0j
    a = abs(rm)
    if abs(cm) != a:
        return 0j
    s = 1.0 / sqrt(2.0)
    par = -1.0 if a % 2 else 1.0
    if rm > 0:
        return complex(s * par) if cm == a else complex(s)
    return complex(0, -s * par) if cm == a else complex(0, s)


def real_cg(l1, r1, l2, r2, L, R):
    """Coupling coefficient of real harmonics: Y_{L,R} component of
    Y_{l1,r1} (x) Y_{l2,r2}, made real by the phase i^(l1+l2-L)."""
    acc = 0j
    for M in range(-L, L + 1):
        u = np.conj(_real_from_complex(R, M))
        if u == 0:
            continue
        for m1 in range(-l1, l1 + 1):
            a = _real_from_complex(r1, m1)
            if a == 0:
                continue
            m2 = M - m1
            if abs(m2) > l2:
                continue
            b = _real_from_complex(r2, m2)
            if b == 0:
                continue
            cg = clebsch_gordan(l1, m1, l2, m2, L, M)
            if abs(cg) < 1e-14:
                continue
            acc += u * cg * a * b
    acc *= (1j) ** ((l1 + l2 - L) % 4)
    if abs(acc.imag) > 1e-10:
        raise ValueError("non-real coupling coefficient")
    return acc.real


class GraphBuilder:
    """Product graph over channels (l, k): tensors of 2l+1 nodes, couplings
    to new tensors, scalar (l=0) outputs per body order."""

    def __init__(self, lmax, kmax, pruning="q-total", cg_cache=None):
        if pruning not in ("q-total", "same-k"):
            raise ValueError("factor pruning must be q-total or same-k")
        self.lmax, self.kmax, self.pruning = lmax, kmax, pruning
        self.q = [(l, k) for l in range(lmax, -1, -1) for k in range(kmax - 1, -1, -1)]
        self.nodes = 0
        self.tensors = {}
        self.basic = []
        self.products = []
        self._plook = {}
        self.scalars = []
        self.scalar_info = []
        self.body_lmax = [lmax] * 7
        self._cg = cg_cache if cg_cache is not None else {}

    def cg(self, *key):
        if key not in self._cg:
            self._cg[key] = real_cg(*key)
        return self._cg[key]

    def _node(self):
        self.nodes += 1
        return self.nodes - 1

    def basic_tensor(self, qi):
        l, k = self.q[qi]
        key = f"q{k * (self.lmax + 1) + l}_l{l}_k{k}"
        if key not in self.tensors:
            nodes = []
            for m in range(-l, l + 1):
                nodes.append(self._node())
                self.basic.append((k, l, m))
            self.tensors[key] = dict(key=key, l=l, nodes=nodes, zero=False)
        return self.tensors[key]

    def add_all_basics(self):
        for qi in range(len(self.q)):
            self.basic_tensor(qi)

    def _add_product(self, a, b, t, c):
        if abs(c) < EPS:
            return
        if b < a:
            a, b = b, a
        key = (a, b, t)
        if key in self._plook:
            self.products[self._plook[key]][3] += c
            return
        self._plook[key] = len(self.products)
        self.products.append([a, b, t, c])

    def couple(self, left, right, L):
        key = f"({left['key']}x{right['key']})->{L}"
        if key in self.tensors:
            return self.tensors[key]
        out = dict(key=key, l=L, nodes=[], zero=True)
        if not (left["zero"] or right["zero"]):
            local, look = [], {}
            l1, l2 = left["l"], right["l"]
            for m1 in range(-l1, l1 + 1):
                for m2 in range(-l2, l2 + 1):
                    for M in range(-L, L + 1):
                        c = self.cg(l1, m1, l2, m2, L, M)
                        if abs(c) < EPS:
                            continue
                        a, b = left["nodes"][m1 + l1], right["nodes"][m2 + l2]
                        if b < a:
                            a, b = b, a
                        lk = (a, b, M + L)
                        if lk in look:
                            local[look[lk]][3] += c
                        else:
                            look[lk] = len(local)
                            local.append([a, b, M + L, c])
            local = [p for p in local if abs(p[3]) >= EPS]
            if local:
                out["zero"] = False
                out["nodes"] = [self._node() for _ in range(2 * L + 1)]
                for a, b, t, c in local:
                    self._add_product(a, b, out["nodes"][t], c)
        self.tensors[key] = out
        return out

    def _allowed(self, qi, body):
        return self.q[qi][0] <= self.body_lmax[body]

    def _tuples(self, count, body):
        same_k = self.pruning == "same-k"
        out = []

        def rec(pos, first, sel_k, tup):
            if pos == count:
                out.append(tuple(tup))
                return
            for qi in range(first, len(self.q)):
                if not self._allowed(qi, body) or (same_k and sel_k >= 0 and self.q[qi][1] != sel_k):
                    continue
                tup.append(qi)
                rec(pos + 1, qi, sel_k if sel_k >= 0 else self.q[qi][1], tup)
                tup.pop()

        rec(0, 0, -1, [])
        return out

    def _specs(self, body_order):
        specs = []
        lq = [q[0] for q in self.q]
        if body_order >= 2:
            specs += [(2, t, 0) for t in self._tuples(1, 2) if lq[t[0]] == 0]
        if body_order >= 3:
            specs += [(3, t, 0) for t in self._tuples(2, 3) if lq[t[0]] == lq[t[1]]]
        if body_order >= 4:
            for t in self._tuples(3, 4):
                l0, l1, l2 = (lq[x] for x in t)
                if (l0 + l1 + l2) % 2 == 0 and _triangle(l0, l1, l2):
                    specs.append((4, t, l2))
        if body_order >= 5:
            for t in self._tuples(4, 5):
                l0, l1, l2, l3 = (lq[x] for x in t)
                if (l0 + l1 + l2 + l3) % 2:
                    continue
                for L in range(max(abs(l0 - l1), abs(l2 - l3)), min(l0 + l1, l2 + l3) + 1):
                    specs.append((5, t, L))
        if body_order >= 6:
            for t in self._tuples(5, 6):
                l0, l1, l2, l3, l4 = (lq[x] for x in t)
                if (l0 + l1 + l2 + l3 + l4) % 2:
                    continue
                for L in range(max(abs(l0 - l1), abs(l2 - l3)), min(l0 + l1, l2 + l3) + 1):
                    if _triangle(L, l4, L):
                        specs.append((6, t, L))
        return specs

    def _scalar(self, body, t, il):
        ts = [self.basic_tensor(x) for x in t[:1]]
        if body == 2:
            return ts[0]["nodes"][0]
        t1 = self.basic_tensor(t[1])
        if body == 3:
            s = self.couple(ts[0], t1, 0)
            return -1 if s["zero"] else s["nodes"][0]
        t2 = self.basic_tensor(t[2])
        if body == 4:
            pair = self.couple(ts[0], t1, il)
            if pair["zero"]:
                return -1
            s = self.couple(pair, t2, 0)
            return -1 if s["zero"] else s["nodes"][0]
        t3 = self.basic_tensor(t[3])
        left = self.couple(ts[0], t1, il)
        right_pair = self.couple(t2, t3, il)
        if body == 6:
            t4 = self.basic_tensor(t[4])
            if left["zero"] or right_pair["zero"]:
                return -1
            right = self.couple(right_pair, t4, il)
            if right["zero"]:
                return -1
            s = self.couple(left, right, 0)
            return -1 if s["zero"] else s["nodes"][0]
        if left["zero"] or right_pair["zero"]:
            return -1
        s = self.couple(left, right_pair, 0)
        return -1 if s["zero"] else s["nodes"][0]

    def add_scalars(self, body_order, body_lmax=None):
        if body_lmax is not None:
            self.body_lmax = list(body_lmax)
        specs = self._specs(body_order)
        need = set()
        for body, t, _ in specs:
            need.update(t)
        for qi in range(len(self.q)):
            if qi in need:
                self.basic_tensor(qi)
        for body, t, il in specs:
            node = self._scalar(body, t, il)
            if node >= 0:
                self.scalars.append(node)
                q = list(t) + [-1] * (5 - len(t))
                self.scalar_info.append((body, *q, il))
        self.products = [p for p in self.products if abs(p[3]) >= EPS]

    def bind_single_channel(self, angular):
        """Same-k scalars of a ranked graph: one single-channel (kmax = 1)
        angular plan replicated over the channels, highest channel first
        (dev_src/sh_model_init.cpp:936-990 BindSingleChannelPlan, used for
        same-k outer and gate rank graphs at :1506-1540).  Channel c's copy
        maps the plan's basic nodes to this graph's (k = c, l, m) basics and
        appends fresh nodes for the rest in the plan's order; scalars are
        interleaved by channel for each plan scalar, with factor indices
        q -> q kmax + (kmax - 1 - c)."""
        if angular.kmax != 1 or angular.lmax != self.lmax or self.products or self.scalars \
                or self.nodes != len(self.basic):
            raise ValueError("a single-channel plan binds onto a graph holding only its basics")
        nb = len(angular.basic)
        where = {b: i for i, b in enumerate(self.basic)}
        nodes = {}
        for c in range(self.kmax - 1, -1, -1):
            m = [where[(c, l, mm)] for _, l, mm in angular.basic]
            m += [self._node() for _ in range(nb, angular.nodes)]
            nodes[c] = m
            self.products += [[m[a], m[b], m[t], v] for a, b, t, v in angular.products]
        for s, info in zip(angular.scalars, angular.scalar_info):
            for c in range(self.kmax - 1, -1, -1):
                self.scalars.append(nodes[c][s])
                q = [x * self.kmax + (self.kmax - 1 - c) if x >= 0 else x for x in info[1:6]]
                self.scalar_info.append((info[0], *q, info[6]))


def _fmt(x):
    return f"{x:.15e}"


def _flist(v):
    return "{" + ", ".join(_fmt(float(x)) for x in v) + "}"


def _ilist(v):
    return "{" + ", ".join(str(int(x)) for x in v) + "}"


def default_scaling(C, blocks):
    """Default warp slopes/shifts per ordered species pair (sus2_sl_model.cpp:217-226)."""
    P = C * C
    v = np.zeros(2 * P * blocks)
    for b in range(blocks):
        for t in range(C):
            for u in range(C):
                x = 0.5 if C == 1 else (sqrt(t + 1.0) + sqrt(u + 1.0) - 2.0) / (2.0 * sqrt(C) - 2.0)
                v[2 * b * P + t * C + u] = 4.0 + (2.0 - 4.0) * x
                v[(2 * b + 1) * P + t * C + u] = 1.5 + x
    return v


def build_model(lmax=3, kmax=4, species=1, body_order=4, outer_rank=2, outer_pruning="q-total",
                gate=False, gate_rank=2, gate_body_order=None, gate_pruning="q-total", rb_size=10,
                cutoff=5.0, min_dist=1.5, scaling=0.01, scaling_map="K", readout="species",
                seed=0, radial_sigma=0.1, gate_w_sigma=0.1, readout_sigma=0.1, tau_sigma=0.0,
                alpha=1.0 / 12.0, name=None):
    """Generate a Static (gate=False) or coordination-Effect (gate=True) model
    with seeded random coefficients; returns the model text."""
    rng = np.random.default_rng(seed)
    C, L1 = species, lmax + 1
    H, F = L1 * L1, kmax * L1
    gate_body_order = body_order if gate_body_order is None else gate_body_order
    cache = {}
    raw = GraphBuilder(lmax, kmax, outer_pruning, cache)
    raw.add_all_basics()
    if not outer_rank:
        raw.add_scalars(body_order)
    g = None

    def ranked(rank, pruning, body):
        # a same-k rank graph is one single-channel plan bound per channel
        # (dev_src/sh_model_init.cpp:1500-1540)
        r = GraphBuilder(lmax, rank, pruning, cache)
        r.add_all_basics()
        if pruning == "same-k":
            a = GraphBuilder(lmax, 1, "same-k", cache)
            a.add_all_basics()
            a.add_scalars(body)
            r.bind_single_channel(a)
        else:
            r.add_scalars(body)
        return r

    if gate:
        if not outer_rank or not gate_rank:
            raise ValueError("the Effect generator builds ranked gate and outer graphs")
        if not 2 <= gate_body_order <= body_order:
            raise ValueError("the gate body order must be in [2, body_order]")
        g = ranked(gate_rank, gate_pruning, gate_body_order)
    o = None
    if outer_rank:
        o = ranked(outer_rank, outer_pruning, body_order)
    goff = raw.nodes
    ooff = raw.nodes + (g.nodes if g else 0)
    M = ooff + (o.nodes if o else 0)
    prods = [list(p) for p in raw.products]
    gpb = len(prods)
    if g:
        prods += [[a + goff, b + goff, t + goff, c] for a, b, t, c in g.products]
    opb = len(prods)
    if o:
        prods += [[a + ooff, b + ooff, t + ooff, c] for a, b, t, c in o.products]
    energy = o if o else raw
    eoff = ooff if o else 0
    scalars = [s + eoff for s in energy.scalars]
    Q = len(scalars)
    blocks = {"pair": 1, "L": L1, "K": kmax, "LK": F}[scaling_map]

    out = ["MTP"]
    ver = "1.8.0" if readout == "species" else ("1.2.0" if outer_rank else "1.1.0")
    out.append(f"version = {ver}")
    out.append(f"potential_name = {name or f'sus2sh_l{lmax}k{kmax}_b{body_order}'}")
    out.append(f"scaling = {_fmt(scaling)}")
    out.append(f"L = {lmax}")
    out.append(f"scaling_map = {scaling_map}")
    out.append(f"species_count = {C}")
    out.append("potential_tag = SUS2-SH")
    out.append(f"sh_l_max = {lmax}")
    out.append(f"sh_k_max = {kmax}")
    out.append(f"sh_body_order = {body_order}")
    out.append("sh_parity = even")
    out.append("sh_body_l_max = " + _ilist([lmax] * (5 if body_order >= 6 else 4)))
    out.append("radial_basis_type = RBChebyshev_sss")
    out.append(f"\tmin_dist = {_fmt(min_dist)}")
    out.append(f"\tmax_dist = {_fmt(cutoff)}")
    out.append(f"\tradial_basis_size = {rb_size}")
    out.append(f"\tradial_funcs_count = {F}")
    out.append("shift_coeffs = " + _flist(np.zeros(C)))
    out.append("scal_coeffs = " + _flist(default_scaling(C, blocks)))
    out.append("\tradial_coeffs")
    out.append("\t\t0-0")
    tau = 1.0 + tau_sigma * rng.standard_normal((F, C)) if tau_sigma else np.ones((F, C))
    for mu in range(F):
        row = list(radial_sigma * rng.standard_normal(rb_size)) + list(tau[mu])
        out.append("\t\t\t" + _flist(row))
    out.append(f"alpha_moments_count = {M}")
    out.append(f"alpha_index_basic_count = {len(raw.basic)}")
    out.append("alpha_index_basic = {" + ", ".join("{%d, %d, %d}" % b for b in raw.basic) + "}")
    out.append("alpha_index_times_count = 0")
    out.append("alpha_index_times = {}")
    out.append(f"sh_product_count = {len(prods)}")
    out.append("sh_products = {" + ", ".join("{%d, %d, %d, %s}" % (a, b, t, _fmt(c)) for a, b, t, c in prods) + "}")
    out.append(f"alpha_scalar_moments = {Q}")
    out.append("alpha_moment_mapping = " + _ilist(scalars))
    out.append(f"sh_scalar_info_count = {Q}")
    out.append("sh_scalar_info = {" + ", ".join("{" + ", ".join(str(x) for x in s) + "}"
                                                for s in energy.scalar_info) + "}")
    if outer_rank:
        out.append("outer_k_rank1_enabled = true")
        out.append(f"outer_k_rank = {outer_rank}")
        # a same-k outer graph holds no cross-channel product (the reader's
        # check, dev_src/mtpr.cpp:12225; the writer's flag, sh_model_init.cpp:1698)
        out.append("outer_channel_mode = " + ("same-channel" if outer_pruning == "same-k" else "cross-channel"))
        out.append("outer_radial_mode = legacy-plus-mixer")
        out.append(f"readout_mode = {readout}")
        out.append(f"outer_k_rank1_basic_begin = {ooff}")
        out.append(f"outer_k_rank1_basic_count = {outer_rank * H}")
        out.append(f"outer_k_rank1_product_begin = {opb}")
        w = np.zeros((L1, outer_rank, kmax))
        for l in range(L1):  # orthonormal rows per l
            q, _ = np.linalg.qr(rng.standard_normal((kmax, kmax)))
            w[l] = q[:outer_rank]
        out.append(f"outer_k_rank1_weight_count = {w.size}")
        out.append("outer_k_rank1_weights = " + _flist(w.ravel()))
    elif readout == "species":
        out.append(f"readout_mode = {readout}")
    if gate:
        out.append(f"two_layer_gate_k_rank = {gate_rank}")
        out.append(f"two_layer_gate_k_rank_basic_begin = {goff}")
        out.append(f"two_layer_gate_k_rank_basic_count = {gate_rank * H}")
        out.append(f"two_layer_gate_k_rank_product_begin = {gpb}")
        out.append(f"two_layer_gate_k_rank_product_end = {opb}")
        w = np.zeros((L1, gate_rank, kmax))
        for l in range(L1):
            q, _ = np.linalg.qr(rng.standard_normal((kmax, kmax)))
            w[l] = q[:gate_rank]
        out.append(f"two_layer_gate_k_rank_weight_count = {w.size}")
        out.append("two_layer_gate_k_rank_weights = " + _flist(w.ravel()))
        Qg = len(g.scalars)
        out.append("two_layer_gate_enabled = true")
        out.append("two_layer_gate_mode = mu-scalar-full")
        out.append(f"two_layer_gate_body_order_max = {gate_body_order}")
        out.append("two_layer_gate_include_one_body = false")
        out.append("two_layer_gate_site_mode = neighbor")
        out.append(f"two_layer_gate_factor_pruning = {gate_pruning}")
        out.append("two_layer_gate_scale_mode = additive-node")
        out.append("two_layer_gate_radial_mode = shared-radial")
        out.append(f"two_layer_gate_radial_coeff_count = {F * rb_size}")
        out.append("two_layer_gate_radial_coeffs = " + _flist(radial_sigma * rng.standard_normal(F * rb_size)))
        out.append(f"two_layer_gate_additive_coeff_count = {C * F}")
        out.append("two_layer_gate_additive_coeffs = " + _flist(np.full(C * F, alpha)))
        out.append(f"two_layer_gate_type_coeff_count = {C * F}")
        out.append("two_layer_gate_type_coeffs = " + _flist(np.ones(C * F)))
        out.append(f"two_layer_gate_weight_count = {F * Qg}")
        out.append("two_layer_gate_scalar_indices = " + _ilist(range(Qg)))
        out.append("two_layer_gate_scalar_moment_indices = " + _ilist([s + goff for s in g.scalars]))
        out.append("two_layer_gate_scalar_body_orders = " + _ilist([s[0] for s in g.scalar_info]))
        out.append("two_layer_gate_weights = " + _flist(gate_w_sigma * rng.standard_normal(F * Qg)))
        out.append("two_layer_gate_edge_l1_enabled = false")
    out.append("species_coeffs = " + _flist(np.zeros(C) if readout == "species" else np.ones(C)))
    nread = Q * (C if readout == "species" else 1)
    out.append("moment_coeffs = " + _flist(readout_sigma * rng.standard_normal(nread)))
    return "\n".join(out) + "\n"


def write_model(path, **kw):
    text = build_model(**kw)
    with open(path, "w") as fh:
        fh.write(text)
    return path
