"""MI355X-native batched MMFF94/MMFF94s conformer minimizer.

Drop-in for the batched MMFF path of mlxmolkit (``mmff_optimize_metal_multi_mol``,
``mmff_optimize_metal_fused_multi_mol``, ``mmff_minimize_nk`` and the ``run_mmff=True``
stage of ``generate_conformers_nk``). Host code is Python on torch-ROCm tensors; the
compute is hand-written HIP for gfx950 behind the C ABI in include/mmff_hip.h.

Public names are imported lazily so that ``import mlxmolkit_archive_amd`` needs neither
torch nor a GPU.
"""
__version__ = "0.1.0"

_EXPORTS = {
    "MMFFParams": "mmff_params",
    "extract_mmff_params": "mmff_params",
    "mmff_minimize_nk": "minimize",
    "mmff_optimize_metal_multi_mol": "optimizer",
    "mmff_optimize_metal_fused_multi_mol": "optimizer",
    "mmff_optimize_metal_batch": "optimizer",
    "MetalBatchResult": "optimizer",
    "MultiMolResult": "optimizer",
    "mmff_energy_grad": "kernel",
    "mmff_energy_grad_metal": "kernel",
    "mmff_energy_grad_metal_mega": "kernel",
    "pack_params_for_metal": "kernel",
    "pack_multi_mol_for_metal": "kernel",
    "generate_conformers_nk": "pipeline",
    "PipelineResult": "pipeline",
    "ConformerResult": "pipeline",
    "Mol": "mol",
    "mol_from_smiles": "mol",
}


def __getattr__(name):
    mod = _EXPORTS.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib
    return getattr(importlib.import_module(f"{__name__}.{mod}"), name)


__all__ = sorted(_EXPORTS)
