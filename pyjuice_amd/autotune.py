"""Measured choices of a circuit's block size, of its chain runs and of each sum layer's kernel
launch.

A structure builder's ``block_size`` only regroups the same nodes (an HCLT with L latents is the
same L x L sum layers at any block size dividing L), but it changes every kernel's tiling: the
number of node blocks per layer, which tile shapes and fast paths apply, and how many
workgroups a narrow layer launches.  The best value depends on the circuit, the batch and the
precision, so it is measured: each candidate is built, compiled and run for a few recorded
training steps (``TrainStepGraph``: clear flows, forward, flow backward, EM), with the chain runs
of ``pyjuice_amd.model.chain`` off and on, and the fastest pair is kept.  The reference fixes ``block_size = num_latents`` for HCLT unless the caller passes one
(``structures/compilation.py:39-40``).

Per layer (``tune_launches``), the launch form of each sum layer's forward and fused element-flow
pass is measured on the circuit's own data: the 16-byte-load kernels with the first two reduction
stages preloaded or not (the default switches at 256 workgroups), the persistent work-queue
forward (``pjamd_sum_forward_queue``) for uneven node blocks, the split reduction off or with
2, 4, 8 or 16 splits, and for FP32 operands the matrix-core or the vector-ALU tiles.
The reference tunes nothing at run time; its tile sizes are fixed per launch
(``layer/sum_layer.py:832-840``).  Choices are cached per layer shape, batch and precision.
"""
from __future__ import annotations

import gc
from typing import Callable, Dict, Iterable, List, Optional, Tuple

import torch

__all__ = ["block_size_candidates", "tune_build", "tune_block_size", "tune_chains", "tune_launches",
           "launch_summary"]


def block_size_candidates(num_latents: int, smallest: int = 64) -> List[int]:
    """Powers of two dividing ``num_latents``, from ``num_latents`` down to ``smallest``."""
    out, bs = [], num_latents
    while bs >= smallest and bs >= 1:
        if num_latents % bs == 0 and bs & (bs - 1) == 0:
            out.append(bs)
        bs //= 2
    return out or [num_latents]


def _time_steps(step: Callable[[], object], steps: int, warmup: int) -> float:
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        step()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def tune_build(make_nodes: Callable[[int], object], x: torch.Tensor, candidates: Iterable[int],
               chain_options: Iterable[bool] = (False,), steps: int = 5, warmup: int = 2,
               precision: Optional[str] = None, lr: float = 0.1, pseudocount: float = 0.1,
               margin: float = 0.03, compile_kwargs: Optional[dict] = None
               ) -> Tuple[Tuple[int, bool], Dict[Tuple[int, bool], float]]:
    """Measured choice of the block size and of the chain runs (``TensorCircuit.use_chains``):
    ``make_nodes(block_size)`` builds the root nodes of one candidate; ``x [B, V]`` (on the
    device) is the batch the recorded training steps run on.  Returns the chosen
    ``(block_size, use_chains)`` and the milliseconds per step of every pair.  A pair replaces
    an earlier one only when faster by more than ``margin`` (relative), so timing noise does not
    flip the choice away from the first (the builder's default block size, chains off).
    ``compile_kwargs`` go to ``juice.compile`` (e.g. ``pflow_dtype``)."""
    import pyjuice_amd as juice
    from pyjuice_amd.optim import TrainStepGraph

    times: Dict[Tuple[int, bool], float] = dict()
    chain_options = list(chain_options)
    for bs in candidates:
        pc = juice.compile(make_nodes(bs), verbose=False, **(compile_kwargs or {}))
        pc.to(x.device)
        for chains in chain_options:
            pc.use_chains = bool(chains)
            g = TrainStepGraph(pc, x, lr=lr, pseudocount=pseudocount, precision=precision)
            times[(bs, bool(chains))] = _time_steps(g, steps, warmup)
            del g
        del pc
        gc.collect()
        torch.cuda.empty_cache()
    best = None
    for k, t in times.items():
        if best is None or t < times[best] * (1.0 - margin):
            best = k
    return best, times


def tune_block_size(make_nodes: Callable[[int], object], x: torch.Tensor, candidates: Iterable[int],
                    steps: int = 5, warmup: int = 2, precision: Optional[str] = None, lr: float = 0.1,
                    pseudocount: float = 0.1, margin: float = 0.03) -> Tuple[int, Dict[int, float]]:
    """``tune_build`` over block sizes alone (layer by layer, no chain runs): the chosen block
    size and the milliseconds per step of every candidate."""
    best, times = tune_build(make_nodes, x, candidates, (False,), steps=steps, warmup=warmup, precision=precision,
                             lr=lr, pseudocount=pseudocount, margin=margin)
    return best[0], {bs: t for (bs, _c), t in times.items()}


def tune_chains(pc, x: torch.Tensor, precision: Optional[str] = None, steps: int = 5, warmup: int = 2,
                lr: float = 0.1, pseudocount: float = 0.1, margin: float = 0.03) -> Dict[bool, float]:
    """Measured choice of ``pc.use_chains`` (recorded training steps on ``x``, chains off first,
    kept unless chains are faster by more than ``margin``); returns the milliseconds per step of
    both settings."""
    from pyjuice_amd.optim import TrainStepGraph

    times: Dict[bool, float] = dict()
    for chains in (False, True):
        pc.use_chains = chains
        g = TrainStepGraph(pc, x, lr=lr, pseudocount=pseudocount, precision=precision)
        times[chains] = _time_steps(g, steps, warmup)
        del g
    pc.use_chains = times[True] < times[False] * (1.0 - margin)
    return times


# (pass, layer shape, batch, precision) -> chosen launch form (None = the library's default rule)
_LAUNCH_CACHE: Dict[tuple, Optional[str]] = dict()


def _layer_key(layer, B: int, prec: str) -> tuple:
    fw = tuple((int(t[0].numel()), int(t[1].size(1)), int(t[5])) for t in layer._dev["fw"])
    bk = tuple((int(t[0].numel()), int(t[1].size(1)), int(t[4]), int(t[5])) for t in layer._dev["bk"])
    return (layer.block_size, fw, bk, tuple(layer._fw_child_run), tuple(layer._fw_aligned),
            tuple(layer._bk_aligned), int(B), prec)


def _fw_candidates(layer, B: int, prec: str) -> List[Optional[str]]:
    from pyjuice_amd.layer.sum_layer import PARAMS_ALIGNED, PRECISIONS
    p = PRECISIONS[prec]
    out: List[Optional[str]] = [None]
    if p == 0 and layer.block_size >= 32:
        return out + ["valu"]          # FP32 operands: matrix cores or vector ALUs
    if p not in (1, 2) or B % 4 != 0 or layer.block_size < 32:
        return out
    bs = layer.block_size
    for e, (t, run, al) in enumerate(zip(layer._dev["fw"], layer._fw_child_run, layer._fw_aligned)):
        flags = run | (PARAMS_ALIGNED if al else 0)
        if al and bs >= 64 and run >= 128 and run % 128 == 0:
            out += ["pre", "nopre"]
            if bs % 128 == 0:
                out.append("wide")
        if layer.queue_ok(flags, p, B):
            out.append("queue")
        plan = layer._split_plan(e, t[0].numel(), t[4], run, al, p, B)
        if plan is not None:
            out.append("nosplit")
            out += [f"split{n}" for n in (2, 4, 8, 16) if n != plan[0] and n <= -(-int(t[4].max()) // 128)]
    return list(dict.fromkeys(out))


def _bk_candidates(layer, B: int, prec: str, fused: bool) -> List[Optional[str]]:
    from pyjuice_amd.layer.sum_layer import PRECISIONS
    p = PRECISIONS[prec]
    out: List[Optional[str]] = [None]
    if p == 0 and not fused and all(t[4] >= 32 for t in layer._dev["bk"]):
        return out + ["valu"]          # FP32 operands, unfused element flows: matrix cores or vector ALUs
    if not fused or p not in (1, 2) or B % 4 != 0 or layer.block_size < 16:
        return out
    if any(al and cbs >= 64 and layer.block_size >= 128 for (_c, _pa, _pp, _k, cbs, _t), al in
           zip(layer._dev["bk"], layer._bk_aligned)):
        out += ["pre", "nopre"]
    for q, (c, _pa, _pp, npar, cbs, _t) in enumerate(layer._dev["bk"]):
        plan = layer._split_plan_bk(q, c.numel(), npar, cbs, layer._bk_aligned[q], p, B)
        if plan is not None:
            out.append("nosplit")
            out += [f"split{n}" for n in (2, 4, 8, 16) if n != plan[0] and n <= int(npar.max()) * layer.block_size // 128]
    return list(dict.fromkeys(out))


def _time_launch(fn: Callable[[], object], reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _pick(times: Dict[Optional[str], float], margin: float) -> Optional[str]:
    best = None
    for c, t in times.items():      # the default (None) is measured first
        if t < times[best] * (1.0 - margin):
            best = c
    return best


def tune_launches(pc, x: torch.Tensor, precision: Optional[str] = None, reps: int = 10, margin: float = 0.03,
                  use_cache: bool = True) -> Dict[str, Dict[str, int]]:
    """Measure every sum layer's candidate launch forms on batch ``x [B, V]`` (on the device) and
    keep the fastest in ``layer.launch_hints`` (``{"fw": ..., "bk": ...}``; None = the library's
    default rule), a candidate replacing the default only when faster by more than ``margin``.
    Each candidate is timed on the layer's own launch, repeated ``reps`` times over the buffers
    of one forward and flow backward of ``x`` (re-running a layer rewrites the same values).
    Recorded step graphs must be recorded after this.  Returns ``launch_summary(pc)``."""
    prec = precision or pc.precision
    B = x.size(0)
    # per pass and candidate: summed ms of the layers it was measured on, and the default's ms on
    # the same layers
    pc._launch_timings = {"fw": dict(), "bk": dict()}
    for g in pc.inner_layer_groups:
        for layer in g:
            if g.is_sum():
                layer.launch_hints = {"fw": None, "bk": None}
    with torch.no_grad():
        pc(x, precision=prec)
        pc.backward(x, flows_memory=0.0, precision=prec)
        fz = pc._fused if prec != "fp32_valu" else {"tables": dict(), "premod": set()}
        for gid, g in enumerate(pc.inner_layer_groups):
            if not g.is_sum():
                continue
            slab = pc._slab(gid)
            for layer in g:
                key = _layer_key(layer, B, prec)
                for pas, cands in (("fw", _fw_candidates(layer, B, prec)),
                                   ("bk", _bk_candidates(layer, B, prec, id(layer) in fz["tables"]))):
                    ck = (pas,) + key
                    if use_cache and ck in _LAUNCH_CACHE:
                        layer.launch_hints[pas] = _LAUNCH_CACHE[ck]
                        continue
                    if len(cands) == 1:
                        continue
                    if pas == "fw":
                        run = lambda: layer(pc.node_mars, slab, pc.params, precision=prec)
                    elif id(layer) not in fz["tables"]:
                        # unfused: element flows of the layer's children, rewritten in place
                        run = lambda: layer.backward(pc.node_flows, pc.element_flows, pc.node_mars, slab, pc.params,
                                                     param_flows=None, allow_modify_flows=True, precision=prec,
                                                     premodified=True)
                    else:
                        # flows already hold log(flow) - mar after the full backward above, so the
                        # layer's pass reads them as premodified and rewrites its children's flows
                        run = lambda: layer.backward(pc.node_flows, pc.element_flows, pc.node_mars, slab, pc.params,
                                                     param_flows=None, allow_modify_flows=True, precision=prec,
                                                     premodified=True, child_tables=fz["tables"][id(layer)])
                    times: Dict[Optional[str], float] = dict()
                    for c in cands:
                        layer.launch_hints[pas] = c
                        times[c] = _time_launch(run, reps)
                    for c in cands[1:]:
                        acc = pc._launch_timings[pas].setdefault(c, {"layers": 0, "ms": 0.0, "default_ms": 0.0})
                        acc["layers"] += 1
                        acc["ms"] += times[c]
                        acc["default_ms"] += times[None]
                    layer.launch_hints[pas] = _pick(times, margin)
                    _LAUNCH_CACHE[ck] = layer.launch_hints[pas]
    return launch_summary(pc)


def launch_summary(pc) -> Dict[str, Dict[str, int]]:
    """How many sum layers use each launch form, per pass ("default" = the library's rule), and
    the measured milliseconds per candidate of the last ``tune_launches`` (``"measured"``)."""
    out: Dict[str, Dict[str, int]] = {"fw": dict(), "bk": dict(),
                                      "measured": getattr(pc, "_launch_timings", None)}
    for g in pc.inner_layer_groups:
        if not g.is_sum():
            continue
        for layer in g:
            h = layer.__dict__.get("launch_hints") or {}
            for pas in ("fw", "bk"):
                k = h.get(pas) or "default"
                out[pas][k] = out[pas].get(k, 0) + 1
    return out
