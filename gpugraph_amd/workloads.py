"""The five benchmark workloads of this engine, with fixed seeds.

Each builder takes a size override so the parity tests can run the same
construction at reduced size.  Seeds: config k uses graph seed k and value
seed 100 + k.

  1. Erdos-Renyi G(n=4096, M=16384) undirected, BFS from vertex 0.
  2. R-MAT(0.57, 0.19, 0.19) scale 22, edge factor 16, undirected, simple;
     int32 weights uniform [1, 64] (symmetric); source = max-degree vertex.
  3. 2048 x 2048 lattice, bonds kept with p = 0.6, + 0.1% long-range pairs.
  4. R-MAT scale 20 directed, int32 weights [1, 64]; 128 SSSP sources and 64
     PPR seeds drawn among vertices with out-degree > 0.
  5. R-MAT scale 22 directed, f32 weights (k+1)/2^24; c ~ N(0, 1).
"""
from __future__ import annotations

import numpy as np

from . import CSR, gen_erdos_renyi, gen_lattice, gen_rmat

RMAT_ABC = (0.57, 0.19, 0.19)


def max_degree_vertex(csr: CSR) -> int:
    """Highest out-degree vertex, lowest id on ties."""
    return int(np.argmax(csr.degrees()))


def config1_er(n: int = 4096, pairs: int = 16384) -> CSR:
    return gen_erdos_renyi(n, pairs, 1)


def config2_rmat(scale: int = 22, weights: bool = True) -> CSR:
    c = gen_rmat(scale, 16, *RMAT_ABC, undirected=True, seed=2)
    if weights:
        c = c.with_weights("i32", 1, 64, symmetric=True, seed=102)
    c.meta.update(config=2)
    return c


def config3_lattice(side: int = 2048) -> CSR:
    c = gen_lattice(side, 0.6, 0.001, 3)
    c.meta.update(config=3)
    return c


def config4_rmat(scale: int = 20) -> CSR:
    c = gen_rmat(scale, 16, *RMAT_ABC, undirected=False, seed=4)
    c = c.with_weights("i32", 1, 64, symmetric=False, seed=104)
    c.meta.update(config=4)
    return c


def config4_instances(csr: CSR, num_sssp: int = 128, num_ppr: int = 64):
    rng = np.random.default_rng(104)
    cand = np.flatnonzero(csr.degrees() > 0)
    picks = rng.choice(cand, num_sssp + num_ppr, replace=False)
    return picks[:num_sssp].astype(np.int32), picks[num_sssp:].astype(np.int32)


def config5_rmat(scale: int = 22) -> CSR:
    c = gen_rmat(scale, 16, *RMAT_ABC, undirected=False, seed=5)
    c = c.with_weights("f32", symmetric=False, seed=105)
    c.meta.update(config=5)
    return c


def config5_loss_weights(n: int) -> np.ndarray:
    return np.random.default_rng(105).standard_normal(n)
