"""Algorithmic FLOP counts (FMA = 2 flops) of the SUS2-SH EFS path.

Counts are per *directed edge* of one pass (each directed edge once, as the
reference's scatter evaluates it), from the loaded model's dimensions:
H = (L+1)^2, F = K(L+1), B radial basis size, Kw warp blocks, KH = K H.

* ``moments``  (one pass, forward):
  geometry 10 + radial basis Kw (13 + 3B) + contraction/envelope F (2B + 3)
  + type weights 3F + harmonics 5H + moment accumulation 2 KH
* ``force``    (one pass, coordinate reverse):
  geometry 10 + radial basis and r-derivative Kw (13 + 7B)
  + contraction/envelope of R and R' F (4B + 6) + weights 4F
  + harmonics and gradient 20H + adjoint contraction 4 KH + 8H (the channel
  sums sum_k abar R and sum_k abar R' per harmonic, then Y and grad Y once)
  (+ 4 KH for the outer pass of an Effect model: the gate adjoint h_bar)
  + force / virial accumulation 21
* ``step_efs`` is the closed-form whole-step count F_e^E (Effect) / F_e^S
  (Static): F_e^S = 11 + 8H + Kw (10B + 12) + 4FB + 11 KH + F + 3,
  F_e^E = F_e^S + 4FB + 13 KH + 2F.

Per-atom terms (``per_atom``; P_g, P_o product rows of the gate and outer
graphs, R, O ranks, Qg gate scalars, Q readout scalars, each product row
v[out] += c v[a] v[b] = 3 flops forward and 6 in reverse):

* ``pass_forward_gate``: rank contraction 2 R KH + graph 3 P_g + h = W S 2 F Qg
* ``pass_forward_outer``: contraction 2 O KH + graph 3 P_o + readout 2 Q,
  then the reverse: graph 6 P_o + transposed contraction 2 O KH
* ``gather_outer`` (Effect): S_bar = h_bar W 2 F Qg + gate graph reverse
  6 P_g + transposed contraction 2 R KH
* ``step_efs`` per atom: 4 (O + R) KH + 9 (P_o + P_g) + 4 F Qg + 2 Q
  (the survey's per-atom term).

Parameter gradient (forward-over-reverse, per directed edge of one pass):

* ``grad_sweep1`` (moments and their tangent along w):
  geometry + edge weight 30 + radial basis and r-derivative Kw (13 + 7B)
  + F (4B + 6) + harmonics and gradient 20H + grad Y . w 5H + staging 4F
  (+ 2F for the outer pass of an Effect model) + accumulation 6 KH
* ``grad_sweep2`` (parameter sweep): geometry + edge weight 30 + harmonics and
  gradient 20H + 5H + adjoint contractions p, q, q' 6 KH + per warp block
  Kw (20 + 9B) (tanh, Chebyshev values and two xi-derivatives) + per channel
  F (6B + 30) (+10F outer Effect: gate messages) + radial-coefficient
  accumulation 11 F B
* ``step_grad`` is the closed-form whole-step count F_e^G = F_e^E + 16 KH + 8 FB + 8 Kw B + 4F.
"""

PEAK = {  # MI355X published peaks
    "fp32_tflops": 157.3,  # vector FP32 with packed (two-wide) FMA
    "fp32_nonpacked_tflops": 78.6,  # one FP32 FMA per lane per clock (this build emits no packed FP32)
    "fp64_tflops": 78.6,
    "hbm_gbs": 8000.0,
}


def graph_stats(path):
    """Product rows of the gate and outer graphs of a .sus2 file: P_g over
    [two_layer_gate_k_rank_product_begin, _end), P_o over
    [outer_k_rank1_product_begin, sh_product_count) (all rows for a model
    without the outer rank block)."""
    f = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line:
                k, v = line.split("=", 1)
                f[k.strip()] = v.strip()
    total = int(f.get("sh_product_count", 0))
    pg = 0
    if f.get("two_layer_gate_enabled", "false") == "true":
        pg = int(f["two_layer_gate_k_rank_product_end"]) - int(f.get("two_layer_gate_k_rank_product_begin", 0))
    po = total - int(f.get("outer_k_rank1_product_begin", 0))
    return dict(P_g=pg, P_o=po)


def _atom_terms(d, g):
    L, K, B, Kw, H, F, KH = dims_of(d)
    R, O, Qg, Q = d.get("R", 0), d.get("O", 0), d.get("Qg", 0), d.get("Q", 0)
    return dict(gate=2 * R * KH + 3 * g["P_g"] + 2 * F * Qg,
                outer=2 * O * KH + 3 * g["P_o"] + 2 * Q + 6 * g["P_o"] + 2 * O * KH,
                gate_reverse=2 * F * Qg + 6 * g["P_g"] + 2 * R * KH)


def step_efs_per_atom(d, g):
    """Per-atom flops of one EFS step (graphs, contractions, gate, readout)."""
    t = _atom_terms(d, g)
    return t["outer"] + (t["gate"] + t["gate_reverse"] if d.get("gated") else 0)


def dims_of(d):
    L, K, B, Kw = d["L"], d["K"], d["B"], d["Kw"]
    H = (L + 1) ** 2
    F = K * (L + 1)
    return L, K, B, Kw, H, F, K * H


def moments_per_edge(d):
    L, K, B, Kw, H, F, KH = dims_of(d)
    return 10 + Kw * (13 + 3 * B) + F * (2 * B + 3) + 3 * F + 5 * H + 2 * KH


def force_per_edge(d, outer_effect=False):
    L, K, B, Kw, H, F, KH = dims_of(d)
    return (10 + Kw * (13 + 7 * B) + F * (4 * B + 6) + 4 * F + 20 * H + 4 * KH + 8 * H
            + (4 * KH if outer_effect else 0) + 21)


def step_efs_per_edge(d):
    L, K, B, Kw, H, F, KH = dims_of(d)
    s = 11 + 8 * H + Kw * (10 * B + 12) + 4 * F * B + 11 * KH + F + 3
    if d.get("gated"):
        s += 4 * F * B + 13 * KH + 2 * F
    return s


def grad_sweep1_per_edge(d, outer_effect=False):
    L, K, B, Kw, H, F, KH = dims_of(d)
    return 30 + Kw * (13 + 7 * B) + F * (4 * B + 6) + 25 * H + 4 * F + (2 * F if outer_effect else 0) + 6 * KH


def grad_sweep2_per_edge(d, outer_effect=False):
    L, K, B, Kw, H, F, KH = dims_of(d)
    return 30 + 25 * H + 6 * KH + Kw * (20 + 9 * B) + F * (6 * B + 30) + (10 * F if outer_effect else 0) + 11 * F * B


def step_grad_per_edge(d):
    L, K, B, Kw, H, F, KH = dims_of(d)
    return step_efs_per_edge(d) + 16 * KH + 8 * F * B + 8 * Kw * B + 4 * F


def kernel_flops(name, d, nedges, natoms=0, graph=None):
    """Algorithmic flops of one launch of the named kernel over nedges
    directed edges and natoms centres (per-atom graph terms when the model's
    graph_stats are given)."""
    t = _atom_terms(d, graph) if graph else dict(gate=0, outer=0, gate_reverse=0)
    if name.startswith("pass_forward_gate"):
        return moments_per_edge(d) * nedges + t["gate"] * natoms
    if name.startswith("pass_forward_outer"):
        return moments_per_edge(d) * nedges + t["outer"] * natoms
    if name.startswith("gather_outer") and d.get("gated") and graph:
        return t["gate_reverse"] * natoms
    if name.startswith("edge_backward_outer"):
        return force_per_edge(d, outer_effect=bool(d.get("gated"))) * nedges
    if name.startswith("edge_backward_gate"):
        return force_per_edge(d) * nedges
    eff = bool(d.get("gated"))
    if name.startswith("grad_gate_forward"):
        return grad_sweep1_per_edge(d) * nedges
    if name.startswith("grad_outer"):
        return (grad_sweep1_per_edge(d, eff) + grad_sweep2_per_edge(d, eff)) * nedges
    if name.startswith("grad_gate_reverse"):
        return grad_sweep2_per_edge(d) * nedges
    return None
