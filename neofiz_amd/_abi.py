"""ctypes prototypes of the C ABI declared in include/neofiz.h.

`bind(lib, prefix)` attaches argument and result types to every entry point of a loaded
library whose symbols carry `prefix` ("nf_" for the HIP library). The structures mirror the C
structs field for field.
"""
import ctypes as C

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)


class JohnsonCookC(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "a", "b", "n", "c", "m", "reference_strain_rate", "room_temperature",
        "melt_temperature", "specific_heat", "taylor_quinney")]


class MaterialC(C.Structure):
    _fields_ = [("youngs_modulus", C.c_double), ("poisson_ratio", C.c_double),
                ("density", C.c_double), ("yield_stress", C.c_double),
                ("isotropic_hardening", C.c_double), ("kinematic_hardening", C.c_double),
                ("has_johnson_cook", C.c_int32), ("reserved", C.c_int32),
                ("johnson_cook", JohnsonCookC)]


class StateC(C.Structure):
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("dt_previous", C.c_double),
                ("dt0", C.c_double), ("mass_scale", C.c_double), ("steps", C.c_int64),
                ("node_count", C.c_int32), ("element_count", C.c_int32),
                ("points_per_element", C.c_int32), ("eroded_elements", C.c_int32),
                ("inverted_elements", C.c_int32), ("wall_contact_count", C.c_int32),
                ("wall_stuck_count", C.c_int32), ("contact_pairs", C.c_int32),
                ("eroded_energy", C.c_double), ("wall_impulse", C.c_double),
                ("wall_energy_loss", C.c_double), ("wall_friction_dissipation", C.c_double),
                ("contact_energy", C.c_double), ("contact_dissipation", C.c_double),
                ("contact_max_penetration", C.c_double), ("contact_escaped_nodes", C.c_int64)]


class AuditsC(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "axial_momentum", "radial_momentum", "kinetic_energy", "centred_kinetic_energy",
        "strain_energy", "plastic_dissipation", "hourglass_energy", "fracture_dissipation",
        "max_plastic_strain", "yielded_fraction", "max_temperature", "max_damage",
        "damaged_fraction", "failed_fraction")]


class DefectParamsC(C.Structure):
    _fields_ = [("scale", C.c_double), ("modulus", C.c_double), ("reference_volume", C.c_double),
                ("correlation_length", C.c_double), ("seed", C.c_int64)]


NODE_FIELDS = {"ur": 0, "uz": 1, "vr": 2, "vz": 3, "fr": 4, "fz": 5, "mass": 6, "r": 7,
               "z": 8, "on_wall": 9, "fixed_r": 10, "fixed_z": 11}
POINT_FIELDS = {"stress": 0, "back": 1, "plastic_strain": 2, "plastic_work": 3,
                "damage_strength": 4, "damage_softening": 5, "damage_width": 6,
                "temperature": 7, "damage": 8, "damage_failure_strain": 9,
                "damage_fracture_energy": 10}
ELEMENT_FIELDS = {"hg_r": 0, "hg_z": 1, "hg_stiffness": 2, "eroded": 3}

_P = C.c_void_p
_PROTOTYPES = {
    "last_error": (C.c_char_p, []),
    "material_check": (C.c_int, [C.POINTER(MaterialC)]),
    "solver_create": (C.c_int, [C.POINTER(_P), C.c_int32, c_double_p, c_double_p, C.c_int32,
                                c_int32_p, C.c_int32, c_int32_p, C.c_int32,
                                C.POINTER(MaterialC), c_int32_p, C.c_int32, C.c_int32,
                                C.c_int32, C.c_double, C.c_double]),
    "solver_destroy": (C.c_int, [_P]),
    "set_threads": (C.c_int, [_P, C.c_int32]),
    "fix_radial": (C.c_int, [_P, C.c_int32]),
    "fix_axial": (C.c_int, [_P, C.c_int32]),
    "fix_all_axial": (C.c_int, [_P]),
    "set_pressure_ramp": (C.c_int, [_P, C.c_double, C.c_double]),
    "set_relaxation_damping": (C.c_int, [_P, C.c_double, C.c_double]),
    "set_hourglass_coefficient": (C.c_int, [_P, C.c_double]),
    "set_rigid_wall": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double]),
    "set_contact": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double, C.c_double]),
    "set_gravity": (C.c_int, [_P, C.c_double, C.c_double]),
    "contact_set_limits": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "set_damage": (C.c_int, [_P, c_double_p, c_double_p]),
    "set_erosion": (C.c_int, [_P, C.c_double]),
    "set_mass_scaling": (C.c_int, [_P, C.c_double]),
    "set_velocity": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double]),
    "set_uniform_velocity": (C.c_int, [_P, C.c_double, C.c_double]),
    "run": (C.c_int, [_P, C.c_int64]),
    "run_until": (C.c_int, [_P, C.c_double, C.c_int64, C.POINTER(C.c_int64)]),
    "advance": (C.c_int, [_P]),
    "get_node_field": (C.c_int, [_P, C.c_int32, c_double_p]),
    "put_node_field": (C.c_int, [_P, C.c_int32, c_double_p]),
    "get_point_field": (C.c_int, [_P, C.c_int32, c_double_p]),
    "get_element_field": (C.c_int, [_P, C.c_int32, c_double_p]),
    "get_state": (C.c_int, [_P, C.POINTER(StateC)]),
    "get_audits": (C.c_int, [_P, C.POINTER(AuditsC)]),
    "temperature": (C.c_int, [_P, C.c_int32, c_double_p]),
    "plastic_front_radius": (C.c_int, [_P, c_double_p]),
    "most_damaged_element": (C.c_int, [_P, c_int32_p]),
    "bore_pressure": (C.c_int, [_P, c_double_p]),
    "mesh_z": (C.c_int, [_P, C.c_int32, c_double_p]),
    "set_damage_uniform": (C.c_int, [_P, C.c_double, c_double_p]),
    "contact_pairs": (C.c_int, [_P, C.c_int32, c_int32_p, c_int32_p]),
    "contact_surface": (C.c_int, [_P, C.c_int32, c_int32_p, c_int32_p]),
    "enable_sensitivity": (C.c_int, [_P]),
    "get_node_sensitivity": (C.c_int, [_P, C.c_int32, C.c_int32, c_double_p]),
    "get_step_sensitivity": (C.c_int, [_P, c_double_p]),
    "adjoint_enable": (C.c_int, [_P, C.c_int32]),
    "adjoint_gradient": (C.c_int, [_P, C.c_int32, c_double_p, c_double_p]),
    "adjoint_info": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(_P)]),
    "batch_destroy": (C.c_int, [_P]),
    "batch_set_tile": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "batch_get_tile": (C.c_int, [_P, C.c_int32, c_int32_p]),
    "batch_run_until": (C.c_int, [_P, C.c_double, C.c_int64]),
    "batch_scatter": (C.c_int, [_P]),
    "math_eval": (C.c_int, [C.c_int32, c_double_p, c_double_p, c_double_p, C.c_int32]),
    "profile": (C.c_int, [_P, C.c_int32]),
    "point_path": (C.c_int, [C.POINTER(MaterialC), C.c_int32, C.c_int32, C.c_int32, c_double_p,
                             c_double_p, c_double_p]),
    "copy_bandwidth": (C.c_int, [C.c_int64, C.c_int32, c_double_p, c_double_p]),
    "profile_read": (C.c_int, [_P, c_double_p, C.POINTER(C.c_int64)]),
    "defect_create": (C.c_int, [C.POINTER(_P), C.POINTER(DefectParamsC)]),
    "defect_destroy": (C.c_int, [_P]),
    "defect_sample_onto": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.c_int32, c_double_p,
                                     c_double_p, C.c_int32, c_int32_p, C.c_int32, C.c_double,
                                     C.c_int32, c_double_p]),
    "defect_sample_onto_device": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.c_int32, _P,
                                            _P, C.c_int32, _P, C.c_int32, C.c_double, C.c_int32,
                                            _P]),
    "defect_sample_onto_static": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.c_int32, _P,
                                            _P, C.c_int32, _P, C.c_int32, C.c_double, C.c_int32,
                                            _P]),
    "defect_points": (C.c_int, [_P, C.c_int32, C.c_int32, c_double_p, c_double_p, c_double_p,
                                c_double_p]),
    "defect_timing": (C.c_int, [_P, c_double_p]),
    "normal_eval": (C.c_int, [C.c_int32, C.c_int32, c_double_p, c_double_p]),
    "film_capture": (C.c_int, [_P, C.c_int32, c_int32_p, c_double_p, c_double_p, c_double_p]),
}


class Binding:
    """The entry points of one library, looked up under one symbol prefix."""

    def __init__(self, lib, prefix):
        self.lib = lib
        self.prefix = prefix
        for name, (res, args) in _PROTOTYPES.items():
            fn = getattr(lib, prefix + name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def symbols(self):
        return [self.prefix + n for n in _PROTOTYPES]
