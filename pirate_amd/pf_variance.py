"""Analytic white-noise variances of the peak finder's outputs: the reference's
PfAvarApproximation (include/pirate/PfVariance.hpp:84-104, src_lib/PfVariance.cpp:238-366,
bound in src_pybind11/pirate_pybind11_avar.cpp:126-140), computed in closed form by
libpirate_amd.so (pirate_amd/csrc/host/pf_variance.hpp states the definition and the
approximation)."""

import ctypes

import numpy as np

from . import _lib as L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class PfAvarApproximation:
    """PfAvarApproximation(plan, freq_variances): for each tree, tree_variance[itree] is a
    float64 (N, ndm_wt, nprofiles) array (the shape fill_host_weights() takes) and
    per_tf[itree] the (2^pf_rank, ndm_wt, nprofiles) frequency-summed pf-rank bands."""

    def __init__(self, plan, freq_variances):
        fv = np.ascontiguousarray(freq_variances, dtype=np.float64)
        if fv.shape != (plan.nfreq,):
            raise ValueError(f'PfAvarApproximation: freq_variances must have shape ({plan.nfreq},), got {fv.shape}')
        self.plan = plan
        self.nfreq = plan.nfreq
        self.ntrees = plan.ntrees
        self.tree_variance, self.per_tf = [], []
        for itree in range(plan.ntrees):
            tv, ptf = self._one_tree(plan, fv, itree)
            self.tree_variance.append(tv)
            self.per_tf.append(ptf)

    @staticmethod
    def _one_tree(plan, fv, itree):
        t = plan.trees[itree]
        tv = np.zeros((t.nsubbands_N, t.ndm_wt, t.nprofiles), dtype=np.float64)
        ptf = np.zeros((2 ** t.pf_rank, t.ndm_wt, t.nprofiles), dtype=np.float64)
        L.check(L.lib().pdd_plan_pf_variances(plan._h, _dptr(fv), fv.size, itree, _dptr(tv), tv.size,
                                              _dptr(ptf), ptf.size))
        return tv, ptf
