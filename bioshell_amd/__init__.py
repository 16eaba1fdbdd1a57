"""MI355X all-vs-all optimal-superposition cRMSD for BioShell ensembles.

Hand-written HIP kernels for gfx950 behind the C ABI of
``include/bioshell_gpu.h``; this package is the Python host mirror of that ABI.
"""
from .crmsd import (  # noqa: F401
    BioshellGpuError,
    Context,
    Library,
    Method,
    RmsdMatrix,
    Server,
    Ticket,
    Rototranslation,
    check,
    crmsd,
    crmsd_all_vs_all,
    crmsd_all_vs_all_device,
    crmsd_all_vs_all_row_stats_device,
    crmsd_all_vs_all_streamed,
    crmsd_batch,
    crmsd_batch_device,
    default_context,
    device_count,
    generate_ensemble,
    superpose,
    tri_index,
    version,
)
from .ensemble import (  # noqa: F401,E402
    Ensemble,
    EnsembleFit,
    load_ensemble,
    superpose_ensemble,
    superpose_ensemble_device,
)
