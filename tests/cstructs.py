"""ctypes mirrors of the reference's structs (hf/structures.h:15-78), for tests
that read what a config reader left in them.  Field order and types follow the
reference header; ctypes lays them out with the C ABI's alignment."""
import ctypes as C

PARAM_FIELDS = [
    ("dimension", C.c_short), ("poly_degree", C.c_short), ("phs_degree", C.c_short),
    ("cloud_size_multiplier", C.c_short), ("test", C.c_short),
    ("courant_number", C.c_double), ("steady_state_tolerance", C.c_double),
    ("poisson_solver_tolerance", C.c_double),
    ("num_vcycles", C.c_short), ("num_relax", C.c_short), ("num_time_steps", C.c_int),
    ("num_levels", C.c_short), ("write_interval", C.c_int), ("neumann_flag_boundary", C.c_bool),
    ("omega", C.c_float), ("dt", C.c_double), ("iter_momentum", C.c_short), ("iter_timple", C.c_short),
    ("rho", C.c_double), ("mu", C.c_double), ("Re", C.c_double), ("facRe", C.c_double),
    ("facdt", C.c_double), ("nu", C.c_double)]


class Params(C.Structure):
    _fields_ = PARAM_FIELDS

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_PTRS = ["x", "y", "z", "point_index", "x_normal", "y_normal", "z_normal", "boundary_tag", "corner_tag",
         "pow_x", "pow_y", "pow_z", "cloud_index", "prolongation_points", "restriction_points", "Dx", "Dy",
         "Dz", "lap", "lap_Poison", "restr_mat", "prol_mat", "prolongation_points2", "restriction_points2"]


class PointStructure(C.Structure):
    _fields_ = ([("mesh_filename", C.c_char * 50), ("num_nodes", C.c_int), ("num_corners", C.c_int),
                 ("num_boundary_nodes", C.c_int), ("num_elem", C.c_int), ("d_avg", C.c_double),
                 ("num_poly_terms", C.c_short), ("num_cloud_points", C.c_short), ("poly_degree", C.c_short)]
                + [(k, C.c_void_p) for k in _PTRS])
