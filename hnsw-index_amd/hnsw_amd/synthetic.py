"""synthetic::Dataset (src/synthetic.rs:62-175) and the benchmark workload
generators, all computed in the host library (hh_synthetic)."""
import numpy as np

from . import _check, _ptr
from ._lib import lib

# generator kinds of hh_synthetic
GAUSSIAN, CLUSTERED, UNIFORM_CLUSTERS, UNIT_UNIFORM, LOGNORMAL, GAUSS_CLUSTERS = range(6)


class Dataset:
    """A base set and a disjoint query set, flat and row-major."""

    def __init__(self, dim, base, queries, categories=None):
        self._dim = dim
        self._base = base
        self._queries = queries
        self.categories = categories

    @staticmethod
    def _make(kind, base, queries, dim, clusters=0, sigma=0.0, sigma_q=0.0, seed=42,
              threads=16, categories=False):
        b = np.empty((base, dim), np.float32)
        q = np.empty((queries, dim), np.float32)
        c = np.empty(base, np.uint32) if categories else None
        _check(lib.hh_synthetic(kind, base, queries, dim, clusters, sigma, sigma_q, seed,
                                threads, _ptr(b), _ptr(q), _ptr(c)))
        return Dataset(dim, b, q, c)

    @staticmethod
    def gaussian(base, queries, dim, seed):
        """src/synthetic.rs:81-92."""
        return Dataset._make(GAUSSIAN, base, queries, dim, seed=seed)

    @staticmethod
    def clustered(base, queries, dim, clusters, seed):
        """src/synthetic.rs:102-124."""
        return Dataset._make(CLUSTERED, base, queries, dim, clusters=clusters, seed=seed)

    @staticmethod
    def uniform_clusters(base, queries, dim, clusters, sigma, sigma_q, seed, threads=16):
        """Centres uniform in [-1,1]^dim; rows centre + sigma*N(0,1), queries sigma_q."""
        return Dataset._make(UNIFORM_CLUSTERS, base, queries, dim, clusters, sigma, sigma_q,
                             seed, threads)

    @staticmethod
    def unit_uniform(base, queries, dim, seed, threads=16):
        """U(0,1) components, then L2-normalised."""
        return Dataset._make(UNIT_UNIFORM, base, queries, dim, seed=seed, threads=threads)

    @staticmethod
    def lognormal(base, queries, dim, sigma, categories, seed, threads=16):
        """N(0,1) rows scaled by exp(sigma*N(0,1)); one category per row below `categories`."""
        return Dataset._make(LOGNORMAL, base, queries, dim, categories, sigma, 0.0, seed,
                             threads, categories=True)

    @staticmethod
    def gauss_clusters(base, queries, dim, clusters, sigma, seed, threads=16):
        """N(0,1) centres; rows and queries centre + sigma*N(0,1)."""
        return Dataset._make(GAUSS_CLUSTERS, base, queries, dim, clusters, sigma, sigma, seed,
                             threads)

    def dim(self): return self._dim
    def base(self): return self._base
    def queries(self): return self._queries
    def base_len(self): return self._base.shape[0]
    def query_len(self): return self._queries.shape[0]
    def base_vector(self, i): return self._base[i]
    def query(self, i): return self._queries[i]
