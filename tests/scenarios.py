"""Synthetic inputs shared by the parity tests (the builders live in
samos_amd.workloads so the benchmark and the training script use the same
ones)."""
from samos_amd.workloads import *  # noqa: F401,F403
from samos_amd.workloads import (ENSEMBLE_COLUMNS, active_replica_params, c4_ensemble_workload, lj_lattice,  # noqa
                                 load, load_ensemble, min_image, plane_disks, plane_replicas, rel_force_err,
                                 sphere_disks, sphere_replicas, surface_disks)
