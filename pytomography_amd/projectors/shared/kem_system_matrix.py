"""H K for a kernel-EM coefficient image (projectors/shared/kem_system_matrix.py:4-58)."""
from __future__ import annotations

from ..system_matrix import SystemMatrix


class KEMSystemMatrix(SystemMatrix):
    """forward = H K alpha, backward = K^T H^T p, sensitivity K^T H_m^T 1.

    The subset helpers are delegated to the wrapped system matrix.  The
    reference binds ``get_weighting_subset`` to the wrapped matrix's
    ``get_projection_subset`` (kem_system_matrix.py:20), which cannot be called
    with a subset index alone; here it is the wrapped ``get_weighting_subset``.
    """

    def __init__(self, system_matrix, kem_transform):
        self.object_meta = system_matrix.object_meta
        self.proj_meta = system_matrix.proj_meta
        self.system_matrix = system_matrix
        self.kem_transform = kem_transform
        self.kem_transform.configure(system_matrix.object_meta, system_matrix.proj_meta)
        self.set_n_subsets = self.system_matrix.set_n_subsets
        self.get_projection_subset = self.system_matrix.get_projection_subset
        self.get_weighting_subset = self.system_matrix.get_weighting_subset

    @property
    def subset_indices_array(self):
        return self.system_matrix.subset_indices_array

    def compute_normalization_factor(self, subset_idx: int | None = None):
        return self.kem_transform.backward(self.system_matrix.compute_normalization_factor(subset_idx))

    def forward(self, object, subset_idx=None):
        return self.system_matrix.forward(self.kem_transform.forward(object), subset_idx)

    def backward(self, proj, subset_idx=None):
        return self.kem_transform.backward(self.system_matrix.backward(proj, subset_idx))
