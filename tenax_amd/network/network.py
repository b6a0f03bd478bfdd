"""TensorNetwork: the MPS/MPO container used by dmrg() (I/O only).

Mirrors the container part of reference src/tenax/network/network.py:29-511
(add_node / get_tensor / replace_tensor / connect / n_nodes / node_ids); the
reference's graph-contraction engine is not on the accelerated path.
"""

from __future__ import annotations


class TensorNetwork:
    def __init__(self, name: str = ""):
        self.name = name
        self._nodes: dict = {}
        self._edges: list = []

    def add_node(self, node_id, tensor) -> None:
        if node_id in self._nodes:
            raise ValueError(f"Node {node_id!r} already exists")
        self._nodes[node_id] = tensor

    def remove_node(self, node_id):
        t = self._nodes.pop(node_id)
        self._edges = [e for e in self._edges if node_id not in (e[0], e[2])]
        return t

    def replace_tensor(self, node_id, tensor) -> None:
        if node_id not in self._nodes:
            raise KeyError(node_id)
        self._nodes[node_id] = tensor

    def get_tensor(self, node_id):
        if node_id not in self._nodes:
            raise KeyError(f"Node {node_id!r} not found")
        return self._nodes[node_id]

    def connect(self, node_a, label_a, node_b, label_b) -> None:
        ta, tb = self.get_tensor(node_a), self.get_tensor(node_b)
        if label_a not in ta.labels():
            raise ValueError(f"Label {label_a!r} not on node {node_a!r}")
        if label_b not in tb.labels():
            raise ValueError(f"Label {label_b!r} not on node {node_b!r}")
        self._edges.append((node_a, label_a, node_b, label_b))

    def node_ids(self) -> list:
        return list(self._nodes)

    def n_nodes(self) -> int:
        return len(self._nodes)

    def n_edges(self) -> int:
        return len(self._edges)

    def neighbors(self, node_id) -> list:
        out = []
        for a, _, b, _ in self._edges:
            if a == node_id:
                out.append(b)
            elif b == node_id:
                out.append(a)
        return out

    def __repr__(self):
        return f"TensorNetwork(name={self.name!r}, nodes={self.n_nodes()}, edges={self.n_edges()})"
