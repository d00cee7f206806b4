"""Base class of the node DAG (vectors of nodes in blocks of ``block_size``).

Behaviour follows ``pyjuice/nodes/nodes.py``: DFS post-order iteration (``nodes.py:13-64``),
power-of-two block sizes (``nodes.py:83``), parameter tying through ``_source_node``.
Traversals are iterative so that deep circuits (long HMM chains, path-like Chow-Liu
trees) do not hit Python's recursion limit; the visiting order equals the recursive one.
"""
from __future__ import annotations

from collections import deque
from typing import Callable, Optional

__all__ = ["CircuitNodes", "node_iterator", "dfs_postorder"]


def dfs_postorder(root, on_finish: Callable, visited: Optional[set] = None):
    """Iterative DFS: children in ``chs`` order, a node is emitted after all its children.

    Identical order to the recursive walk of ``nodes.py:14-27``.
    """
    if visited is None:
        visited = set()
    if root in visited:
        return
    visited.add(root)
    stack = [[root, 0]]
    while stack:
        top = stack[-1]
        ns, i = top
        chs = ns.chs if (ns.is_sum() or ns.is_prod()) else ()
        if i < len(chs):
            top[1] = i + 1
            c = chs[i]
            if c not in visited:
                visited.add(c)
                stack.append([c, 0])
        else:
            stack.pop()
            on_finish(ns)


def node_iterator(root_ns, reverse: bool = False):
    if not reverse:
        out = []
        dfs_postorder(root_ns, out.append)
        return iter(out)

    # Reverse topological order (parents before children), as nodes.py:37-61.
    parcount = dict()

    def inc(ns):
        for cs in ns.chs:
            parcount[cs] = parcount.get(cs, 0) + 1

    dfs_postorder(root_ns, inc)
    out = []
    q = deque([root_ns])
    while q:
        ns = q.popleft()
        out.append(ns)
        for cs in ns.chs:
            parcount[cs] -= 1
            if parcount[cs] == 0:
                q.append(cs)
    assert len(parcount) + 1 == len(out)
    return iter(out)


class CircuitNodes:
    INIT_CALLBACKS = []
    DEFAULT_BLOCK_SIZE = 1

    def __init__(self, num_node_blocks: int, region_node, block_size: int = 0,
                 source_node: Optional["CircuitNodes"] = None, **kwargs):
        if block_size == 0:
            block_size = CircuitNodes.DEFAULT_BLOCK_SIZE
        assert num_node_blocks > 0
        assert block_size > 0 and (block_size & (block_size - 1)) == 0, \
            f"block_size must be a power of two (got {block_size})"

        self.num_node_blocks = int(num_node_blocks)
        self.block_size = int(block_size)
        self.region_node = region_node
        self.chs = []

        self._output_ind_range = None
        self._param_ids = None
        self._params = None

        if source_node is not None:
            while source_node._source_node is not None:
                source_node = source_node._source_node
        self._source_node = source_node
        self._tied_param_block_ids = None
        self._reverse_iter = False

    def _run_init_callbacks(self, **kwargs):
        for fn in CircuitNodes.INIT_CALLBACKS:
            fn(self, **kwargs)

    @property
    def scope(self):
        return self.region_node.scope

    def is_sum(self):
        return False

    def is_prod(self):
        return False

    def is_input(self):
        return False

    @property
    def num_chs(self):
        return len(self.chs)

    @property
    def num_nodes(self):
        return self.num_node_blocks * self.block_size

    @property
    def num_edges(self):
        raise NotImplementedError()

    @property
    def num_vars(self):
        return len(self.scope)

    def duplicate(self, *args, **kwargs):
        raise ValueError(f"{type(self).__name__} cannot be duplicated")

    def _draw_own_params(self, perturbation: float, **kwargs):
        """Random initialisation of this vector's own parameters (no recursion)."""
        return None

    def init_parameters(self, perturbation: float = 2.0, recursive: bool = True, is_root: bool = True, **kwargs):
        """Randomly initialise parameters.

        Draw order follows ``nodes.py:124-141`` with the subclass overrides: a sum vector
        draws its own parameters on every visit, then the walk descends into children it
        has not visited yet (pre-order); input vectors draw only once.
        """
        self._draw_own_params(perturbation, **kwargs)
        visited = {self}
        if recursive:
            stack = [iter(self.chs)]
            while stack:
                try:
                    cs = next(stack[-1])
                except StopIteration:
                    stack.pop()
                    continue
                cs._draw_own_params(perturbation, **kwargs)
                if cs in visited:
                    continue
                visited.add(cs)
                stack.append(iter(cs.chs))
        for ns in visited:
            if ns._source_node is not None and not ns.is_input():
                ns._params = None

    def num_parameters(self, count_input_params: bool = False, recursive: bool = True):
        nodes = list(self) if recursive else [self]
        cnt = 0
        for ns in nodes:
            if ns.is_sum() and not ns.is_tied():
                cnt += ns.num_edges
            elif ns.is_input() and count_input_params:
                cnt += ns.num_nodes * ns.dist.num_parameters()
        return cnt

    def is_tied(self):
        return self._source_node is not None

    def get_params(self, *args, **kwargs):
        return None

    def get_param_flows(self, *args, **kwargs):
        return None

    def get_source_ns(self):
        return self._source_node if self.is_tied() else self

    def set_source_ns(self, source_ns: "CircuitNodes"):
        assert type(source_ns) == type(self)
        assert len(source_ns.chs) == len(self.chs)
        assert getattr(self, "_params", None) is None
        assert source_ns.num_node_blocks == self.num_node_blocks
        assert source_ns.block_size == self.block_size
        self._source_node = source_ns

    def has_params(self):
        if self.is_input():
            return self._param_initialized
        src = self.get_source_ns()
        return getattr(src, "_params", None) is not None

    def contains(self, ns: "CircuitNodes"):
        return any(n is ns for n in self)

    def _clear_tensor_circuit_hooks(self, recursive: bool = True):
        for ns in (list(self) if recursive else [self]):
            for name in ("_param_range", "_param_ids", "_inverse_param_ids", "_param_flow_range",
                         "_output_ind_range", "_param_flow_ids"):
                if hasattr(ns, name):
                    setattr(ns, name, None)

    def __len__(self):
        n = [0]

        def cnt(_):
            n[0] += 1

        dfs_postorder(self, cnt)
        return n[0]

    def __iter__(self):
        return node_iterator(self, self._reverse_iter)

    def __call__(self, reverse: bool = False):
        self._reverse_iter = reverse
        return self

    def provided(self, var_name):
        return getattr(self, var_name, None) is not None

    # Identity semantics for hashing (CircuitNodes are used as dict keys everywhere).
    __hash__ = object.__hash__
