"""momentumed — exact diagonalisation of momentum-conserved fermion systems on MI355X.

Drop-in mirror of the MomentumED.jl API for the matrix-free device path:
    para = EDPara(k_list=..., Gk=..., H_two=...)
    subspaces, ks = ED_momentum_subspaces(para, N)          # enumerated on the GPU
    H = MBOperator(ED_scatterlist_twobody(para), upper_hermitian=True)
    energies, vectors = EDsolve(subspaces[0], H, N=6, method="map", device="gpu")
The compute runs in libedm.so (hand-written gfx950 HIP kernels behind a C-ABI).
"""
from .core import (ColexMBS64, ColexMBS64Mask, HilbertSubspace, MBOperator, MBS64,
                   MBS64_complete, MBS64Vector, Scatter, empty, flip, for_each_occ,
                   get_bits, get_body,
                   index_fit, isdiagonal, isempty, isoccupied, isphysical, isupper,
                   make_mask64, occ_list, occupy, parity_mask, scat_occ_number,
                   sort_merge_scatlist)
from .device import (DEFAULT_GPU, GPU_AVAILABLE, GPU_MEMORY_MONITOR, LinearMap,
                     SparseHmltMatrix, activate_gpu_device, create_gpu_linearmap,
                     create_gpu_matrix, krylov_map_solve, krylov_matrix_solve,
                     release_gpu_memory)
from .preparation import (EDPara, ED_momentum_subspaces, ED_scatterlist_onebody,
                          ED_scatterlist_twobody, ED_sortedScatterList_onebody,
                          ED_sortedScatterList_twobody, MBS_totalmomentum, momentum_residue)
from .solve import EDsolve
from .batch import BatchSolver, EDsolve_batch
from .analysis import (ED_bracket, ED_bracket_threaded, ED_connection_gaugefixing,
                       ED_step_inner_prod, PES_1rdm, PES_MomtBlock_rdm)

__all__ = [
    "EDPara", "ED_momentum_subspaces", "ED_scatterlist_onebody", "ED_scatterlist_twobody",
    "ED_sortedScatterList_onebody", "ED_sortedScatterList_twobody", "MBS_totalmomentum",
    "momentum_residue", "HilbertSubspace", "MBOperator", "MBS64", "MBS64Vector", "Scatter",
    "get_body", "index_fit", "isdiagonal", "isupper", "make_mask64", "occ_list", "parity_mask",
    "sort_merge_scatlist", "EDsolve", "GPU_AVAILABLE", "DEFAULT_GPU", "activate_gpu_device",
    "create_gpu_linearmap", "krylov_map_solve", "release_gpu_memory", "BatchSolver",
    "create_gpu_matrix", "krylov_matrix_solve", "GPU_MEMORY_MONITOR", "LinearMap",
    "SparseHmltMatrix", "EDsolve_batch", "ED_bracket", "ED_bracket_threaded",
    "ED_connection_gaugefixing", "ED_step_inner_prod", "PES_1rdm", "PES_MomtBlock_rdm",
    "ColexMBS64", "ColexMBS64Mask", "MBS64_complete", "empty", "flip", "get_bits",
    "isempty", "isoccupied", "isphysical", "occupy", "scat_occ_number", "for_each_occ",
]
