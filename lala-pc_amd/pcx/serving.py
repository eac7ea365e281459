"""Seeded serving traces: Poisson arrivals of propagation requests.

A trace holds a pool of seed-generated networks whose sizes are drawn
log-uniformly between two propagator counts (variables = propagators / 4),
and a stream of requests.  Request i arrives after exponential gaps of mean
1/rate seconds, picks one network of the pool uniformly, and carries that
network's store with 1..max_splits random splits (x <= v or x > v, v inside
the current bounds).  Everything here is host code and deterministic in the
seed; the networks are registered once with a Server and stay resident in
device memory, as the request stream reuses them.
"""
import numpy as np

from . import gen_network, gen_params, gen_splits


class Trace:
    """nets: list of (op, x, y, z, lb, ub) with lb/ub already embedded by tell;
    arrival_s, net, store_off, lb_all, ub_all: the request stream."""

    def __init__(self, nets, arrival_s, net, store_off, lb_all, ub_all):
        self.nets = nets
        self.arrival_s = arrival_s
        self.net = net
        self.store_off = store_off
        self.lb_all = lb_all
        self.ub_all = ub_all

    def __len__(self):
        return len(self.arrival_s)

    def store(self, i):
        s = slice(int(self.store_off[i]), int(self.store_off[i + 1]))
        return self.lb_all[s], self.ub_all[s]

    def nprops(self, i):
        return len(self.nets[int(self.net[i])][0])


def make_trace(seed, n_req, rate=2000.0, n_nets=24, pmin=1000, pmax=1000000, min_splits=1,
               max_splits=10, embed=None):
    """Build a trace.  embed(op, x, y, z, lb, ub) applies the engine's tell-time
    embedding of reified outputs to a network's store (PIR.tell does it)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(pmin), np.log(pmax), n_nets)).astype(np.int64)
    nets = []
    for j, p_j in enumerate(sizes):
        v_j = max(16, int(p_j) // 4)
        p = gen_params(seed=seed * 1000 + j, nvars=v_j, nprops=int(p_j))
        op, x, y, z, lb, ub, _ = gen_network(p)
        if embed is not None:
            embed(op, x, y, z, lb, ub)
        nets.append((op, x, y, z, lb, ub))
    gaps = rng.exponential(1.0 / rate, n_req)
    arrival = np.cumsum(gaps) - gaps[0]
    net = rng.integers(0, n_nets, n_req).astype(np.int32)
    nv = np.array([len(nets[k][4]) for k in net], np.int64)
    store_off = np.zeros(n_req + 1, np.int64)
    np.cumsum(nv, out=store_off[1:])
    lb_all = np.empty(int(store_off[-1]), np.int32)
    ub_all = np.empty(int(store_off[-1]), np.int32)
    for j in range(n_nets):
        reqs = np.nonzero(net == j)[0]
        if len(reqs) == 0:
            continue
        lb, ub = nets[j][4], nets[j][5]
        lo, hi = gen_splits(seed * 1000 + 500 + j, lb, ub, len(reqs), min_splits, max_splits)
        lo, hi = lo.reshape(len(reqs), -1), hi.reshape(len(reqs), -1)
        for r, i in enumerate(reqs):
            s = slice(int(store_off[i]), int(store_off[i + 1]))
            lb_all[s] = lo[r]
            ub_all[s] = hi[r]
    return Trace(nets, arrival.astype(np.float64), net, store_off, lb_all, ub_all)
