"""Writes tests/golden/checkpoint_v590_p2.txt: a checkpoint in the layout the
reference reads for version 590 -- the oldest branch of its unpack methods
(src/Parameters.cc:476-586: six deprecated solver keys for versions < 592,
init_file_name / cntrs_file_name for < 593, no name line and no
vesicle_props_file line for <= 590; src/EvolveSurface.cc:494-521: no
VESICLEPROPS block for <= 590; src/Vectors.cc:100-127, src/Scalars.cc:238-283:
a name line for <= 590).  The reference repository holds no checkpoint
file, so this fixture is assembled from those branches: one vesicle at
p = 2 whose 36 position values are 0.25 k - 4, k = 0..35.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    lines = ["PARAMETERS", "version: 590", "n_surfs: 1", "sh_order: 2", "filter_freq: 1",
             "upsample_freq: 4", "bending_modulus: 2.5e-02", "viscosity_contrast: 1",
             "position_solver_iter: 15", "tension_solver_iter: 15",
             "position_solver_restart: 1", "tension_solver_restart: 1",
             "position_solver_tol: 1e-06", "tension_solver_tol: 1e-06",
             "time_horizon: 0.5", "ts: 0.001", "time_tol: 1e-06", "time_iter_max: 100",
             "time_adaptive: 0", "solve_for_velocity: 1", "pseudospectral: 0",
             "scheme: GloballyImplicit", "time_precond: DiagonalSpectral", "bg_flow: ShearFlow",
             "singular_stokes: ViaSpHarm", "rep_type: PolyKReparam", "rep_maxit: 10",
             "rep_upsample: 0", "rep_filter_freq: 2", "rep_ts: 0.001", "rep_tol: 0.0001",
             "rep_exponent: 4", "repul_dist: 0.05", "bg_flow_param: 0.1", "periodic_length: -1",
             "interaction_upsample: 0", "checkpoint: 1", "checkpoint_stride: 0.01",
             "write_vtk:  |", "init_file_name: precomputed/shape_gallery_2.txt |",
             "cntrs_file_name:  |", "checkpoint_file_name: run_{{time_idx}}.chk |",
             "load_checkpoint:  |", "error_factor: 1", "num_threads: 4", "excess_density: 0",
             "gravity_field: 0 0 -1", "/PARAMETERS",
             "EVOLVE", "version: 590", "name: evolve_surface",
             "SURFACE", "version: 590", "name: surface", "sh_order: 2", "SHT_order: 2",
             "SHT_rep_order: 2", "SHT_rep_exponent: 4", "SHT_resample_order: 4",
             "diff_filter_freq: 1", "reparam_filter_freq: 2", "reparam_type: PolyKReparam",
             "VECTORS", "version: 590", "name: position", "CoordinateOrder: AxisMajor",
             "SCALARS", "version: 590", "name: position", "nsubs: 1", "shorder: 2", "gridx: 3",
             "gridy: 4", "stride: 12",
             "ARRAY", "version: 590", "name: position", "size: 36",
             "data: " + " ".join("%.17g" % (0.25 * k - 4) for k in range(36)),
             "/ARRAY", "/SCALARS", "/VECTORS", "/SURFACE", "/EVOLVE"]
    with open(os.path.join(HERE, "checkpoint_v590_p2.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
