"""ctypes mirror of include/mpsr.h.

The structs follow the reference PODs field for field (solver/Common/inc/DataType.h:296-627);
see include/mpsr.h for the per-struct citations.
"""
import ctypes as C

import numpy as np

MPSR_NOERR, MPSR_FILEERR, MPSR_PARAMERR, MPSR_CALCERR, MPSR_MEMERR, MPSR_GPUERR, MPSR_ETCERR = range(7)
STATUS_NAMES = ["NOERR", "FILEERR", "PARAMERR", "CALCERR", "MEMERR", "GPUERR", "ETCERR"]
MPSR_FP32, MPSR_FP64 = 0, 1
MPSR_PRESSURE_EOS, MPSR_PRESSURE_PPE = 0, 1
MPSR_PCG_FP64, MPSR_PCG_MIXED = 0, 1
MPSR_WARM_ZERO, MPSR_WARM_PREVIOUS, MPSR_WARM_LEARNED = 0, 1, 2


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


_PARAM_FIELDS = [
    ("Dimension", C.c_int), ("GravityConstant", Float3), ("Density", C.c_float),
    ("ConstantParticleDensity", C.c_float), ("SurfaceTensionCoefficient", C.c_float),
    ("InitialParticleDistance", C.c_float), ("FreeSurfaceThreshold", C.c_float),
    ("GradientInfluenceRadiusCoefficient", C.c_float), ("LaplacianInfluenceRadiusCoefficient", C.c_float),
    ("TemperatureInfluenceRadiusCoefficient", C.c_float), ("FruidInitializeialTemperature", C.c_float),
    ("WallInitializeialTemperature", C.c_float), ("FruidThermalResistance", C.c_float),
    ("WallThermalResistance", C.c_float), ("FruidThermalConductivity", C.c_float),
    ("WallThermalConductivity", C.c_float), ("FruidSpecificHeat", C.c_float), ("FruidDensity", C.c_float),
    ("Dt", C.c_float), ("AnalysisStep", C.c_int), ("ThermalStep", C.c_int), ("NegativeCalculationStep", C.c_int),
    ("OutputStep", C.c_int), ("LambdaID", C.c_int), ("SORJacobianRelaxationCoefficient", C.c_float),
    ("SORLoopMax", C.c_int), ("SORConvergenceConstant", C.c_float), ("bUpdateTemperature", C.c_int),
    ("LambdaValue", C.c_float), ("CellSize", Float3), ("MvVelocity", Float3), ("CourantNumber", C.c_float),
    ("ViscosityUpdateParameterA", C.c_float), ("ViscosityUpdateParameterB", C.c_float),
    ("MaxWallTemperature", C.c_float), ("MinWallTemperature", C.c_float), ("OffSet", C.c_float),
    ("ParticleGenerationStep", C.c_int), ("TotalGenerationStep", C.c_int), ("Alpha", C.c_float),
    ("AlphaLightWeight", C.c_float), ("MaxNeighborhoodNum", C.c_int), ("bExplicitCollisonCheck", C.c_int),
    ("bExplicitCollisionOut", C.c_int), ("bImplicitCollisonCheck", C.c_int),
    ("CollisionCheckCoefficient", C.c_float), ("MinDistanceCoefficient", C.c_float),
    ("bFreeSurfacePressureUpdate", C.c_int), ("BasePressure", C.c_float), ("MaxPressure", C.c_float),
    ("Time1", C.c_float), ("TriggerStep", C.c_int), ("EndStep", C.c_int), ("AnalysisArea", C.c_float),
    ("bExplicitOut", C.c_int), ("WallFrictionCoeffieicnt", C.c_float), ("DragCoefficient", C.c_float),
    ("bTwoPhase", C.c_int), ("StopWallStep", C.c_int), ("Tolerance", C.c_float),
    ("ImplicitSurfaceThreshold", C.c_float), ("bCGSolver", C.c_int), ("bUpdateInnerPressure", C.c_int),
    ("SurfaceInfluenceRadiusCoefficient", C.c_float), ("SurfaceThreshold", C.c_float),
    ("GradientForce", C.c_float), ("StartStep", C.c_int), ("InitialForce", C.c_float),
    ("VelocityThreshold", C.c_float), ("bExplicitPressure", C.c_int), ("SonicVelocity", C.c_float),
    ("ArtificialPressure", C.c_float), ("r_coeff", C.c_float), ("bImplicitVelocity", C.c_int),
    ("bNJPCGSolver", C.c_int), ("PressureMax", C.c_float), ("RLimitCoefficient", C.c_float),
    ("bTurbulance", C.c_int), ("SmagorinskyConstant", C.c_float), ("FilterWidth", C.c_float),
    ("InitialTimeForZeroAngle", C.c_float), ("FinalTimeToReachAngle", C.c_float),
    ("DragPermiabilityConstant", C.c_float),
]


class Params(C.Structure):
    _fields_ = _PARAM_FIELDS


class Box(C.Structure):
    _fields_ = [("m_MinBound", Float3), ("m_MaxBound", Float3)]


class GridBox(C.Structure):
    _fields_ = [("m_BufferedZone", Box), ("m_ComputeZone", Box), ("m_PressureZone", Box),
                ("m_ParentZone", Box), ("m_GridSize", Int3), ("m_CellSize", C.c_float)]

    def cells(self):
        g = self.m_GridSize
        return g.x * g.y * g.z


class Triangle(C.Structure):
    _fields_ = [("ModelIDNumber", C.c_int), ("Normal", Float3), ("Vertex", Float3 * 3)]


class TriParams(C.Structure):
    _fields_ = [("Velocity", Float3), ("Temperature", C.c_float), ("LinearMovableOrNot", C.c_int),
                ("LinearResetInterval", C.c_int), ("RotatingOrNot", C.c_int),
                ("AngleOfRotationInDegree", C.c_float), ("RotationStartTime", C.c_float),
                ("TimeToReachRPM", C.c_float), ("CenterOfRotation", Float3),
                ("SecondPointOfRotation", Float3), ("WallFrictionCoefficient", C.c_float),
                ("WallThermalResistivity", C.c_float), ("WallThermalConductivity", C.c_float),
                ("isDrag", C.c_int), ("DragTriangleID", C.c_int)]


class DragParams(C.Structure):
    _fields_ = [("DragID", C.c_int), ("DragToken", C.c_char), ("InputVelocityToken1", C.c_char),
                ("OutputVelocityToken2", C.c_char), ("CoefficientofInflueneRegion", C.c_float),
                ("VelocityIndex", C.c_float), ("VelocityExponential", C.c_float),
                ("VelocityDragPermiabilityConstant", C.c_float), ("TempIndex", C.c_float),
                ("TempExponential", C.c_float), ("TempDragPermiabilityConstant", C.c_float),
                ("ConstantVx", C.c_float), ("ConstantVy", C.c_float), ("ConstantVz", C.c_float),
                ("VelocityMagnitueFactor", C.c_float)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class ParticlesC(C.Structure):
    _fields_ = [("n", C.c_int), ("id", _ip), ("position", _dp), ("velocity", _dp), ("pressure", _dp),
                ("density", _dp), ("temperature", _dp), ("kinetic_viscosity", _dp),
                ("solid_phase_rate", _dp), ("type", _ip)]


MPSR_PRECOND_JACOBI, MPSR_PRECOND_TWO_LEVEL, MPSR_PRECOND_CELL_VCYCLE = 0, 1, 2


class PpeOptions(C.Structure):
    _fields_ = [("solver", C.c_int), ("rel_tol", C.c_double), ("max_iter", C.c_int),
                ("warm_start", C.c_int), ("n0_lap", C.c_double), ("lambda_lap", C.c_double),
                ("precond", C.c_int), ("coarse_cells", C.c_int), ("coarse_rebuild", C.c_int),
                ("learned_omega", C.c_int), ("source_relaxation", C.c_double)]


class RadixTerms(C.Structure):
    """mpsr_radix_terms: the radix variant's turbulence and negative-pressure switches
    (radix/SPVSOLVER/MultiGPUSolver/DataType.h:510-520)."""
    _fields_ = [("bTurbulance", C.c_int), ("SmagorinskyConstant", C.c_float), ("FilterWidth", C.c_float),
                ("TurbulanViscosityRatio", C.c_float), ("Cnue", C.c_float), ("Cepsilon", C.c_float),
                ("TurbulanceIntensity", C.c_float), ("bNegativePressure", C.c_int)]


class NegativePressureModel(C.Structure):
    """mpsr_negative_pressure_model: per-model fields of the radix CTriangleParameters
    (DataType.h:160-168)."""
    _fields_ = [("RotationInRPM", C.c_float), ("NegativePressureCheck", C.c_int), ("InfluenceRegion", C.c_float),
                ("DirectionVector", Float3), ("PressureGas", C.c_float)]


def radix_terms_args(terms, models):
    """(RadixTerms, NegativePressureModel array or None, count) from structs or dicts."""
    t = terms if isinstance(terms, RadixTerms) else RadixTerms(**dict(terms))
    if models is None:
        return t, None, 0
    arr = (NegativePressureModel * max(1, len(models)))()
    for i, m in enumerate(models):
        if isinstance(m, NegativePressureModel):
            arr[i] = m
        else:
            m = dict(m)
            dv = m.pop("DirectionVector", (0.0, 0.0, 0.0))
            arr[i] = NegativePressureModel(DirectionVector=Float3(*dv), **m)
    return t, arr, len(models)


class SpheresC(C.Structure):
    _fields_ = [("n", C.c_int), ("position", _dp), ("velocity", _dp), ("force", _dp), ("diameter", C.c_double),
                ("density", C.c_double), ("fluid_viscosity", C.c_double)]


class CreateInfo(C.Structure):
    _fields_ = [("params", Params), ("compute_zone", Box), ("max_particles", C.c_int),
                ("particles", C.POINTER(ParticlesC)), ("triangles", C.POINTER(Triangle)),
                ("num_triangles", C.c_int), ("models", C.POINTER(TriParams)), ("num_models", C.c_int),
                ("drag", C.POINTER(DragParams)), ("num_drag", C.c_int), ("precision", C.c_int),
                ("pressure_mode", C.c_int), ("ppe", PpeOptions), ("device", C.c_int)]


class Timings(C.Structure):
    _fields_ = [("step_ms", C.c_double), ("bin_ms", C.c_double), ("tri_ms", C.c_double),
                ("stl_ms", C.c_double), ("drag_ms", C.c_double), ("explicit_ms", C.c_double),
                ("relocate_ms", C.c_double), ("pnd_ms", C.c_double), ("ppe_ms", C.c_double),
                ("gradient_ms", C.c_double), ("wall_ms", C.c_double), ("ppe_iterations", C.c_int),
                ("ppe_refinements", C.c_int), ("ppe_rel_residual", C.c_double),
                ("ppe_initial_rel_residual", C.c_double), ("spmv_ms", C.c_double),
                ("spmv_launches", C.c_int), ("spmv_bytes", C.c_double), ("ppe_nnz", C.c_longlong),
                ("ppe_rows", C.c_int), ("ppe_fallbacks", C.c_int),
                ("ppe_coarse_rows", C.c_int), ("ppe_coarse_rebuilt", C.c_int), ("ppe_setup_ms", C.c_double),
                ("ppe_learned", C.c_int), ("ppe_learned_fallback", C.c_int), ("sphere_ms", C.c_double),
                ("ppe_stored", C.c_longlong), ("thermal_ms", C.c_double), ("turbulence_ms", C.c_double),
                ("steps", C.c_int), ("ppe_iterations_total", C.c_longlong), ("ppe_rebuilds", C.c_int),
                ("ppe_rel_residual_max", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ParticleArrays:
    """numpy-owned host buffers for one mpsr_particles view (the caller owns host memory)."""

    def __init__(self, n):
        self.n = n
        self.id = np.zeros(n, np.int32)
        self.position = np.zeros((n, 3), np.float64)
        self.velocity = np.zeros((n, 3), np.float64)
        self.pressure = np.zeros(n, np.float64)
        self.density = np.zeros(n, np.float64)
        self.temperature = np.zeros(n, np.float64)
        self.kinetic_viscosity = np.zeros(n, np.float64)
        self.solid_phase_rate = np.zeros(n, np.float64)
        self.type = np.zeros(n, np.int32)

    def struct(self):
        def d(a):
            return a.ctypes.data_as(_dp)

        def i(a):
            return a.ctypes.data_as(_ip)
        return ParticlesC(self.n, i(self.id), d(self.position), d(self.velocity), d(self.pressure),
                          d(self.density), d(self.temperature), d(self.kinetic_viscosity),
                          d(self.solid_phase_rate), i(self.type))

    def copy(self):
        o = ParticleArrays(self.n)
        for k in ("id", "position", "velocity", "pressure", "density", "temperature",
                  "kinetic_viscosity", "solid_phase_rate", "type"):
            setattr(o, k, getattr(self, k).copy())
        return o

    FIELDS = ("id", "position", "velocity", "pressure", "density", "temperature",
              "kinetic_viscosity", "solid_phase_rate", "type")


def triangles_array(tris):
    arr = (Triangle * max(1, len(tris)))()
    for i, t in enumerate(tris):
        arr[i] = t
    return arr


class DtEntry(C.Structure):
    """mpsr_dt_entry: one record of the time-step schedule (DeltaTime, Source/DtFrame.cpp:11-35)."""
    _fields_ = [("step", C.c_int), ("dt", C.c_float), ("courant", C.c_float), ("inflow_step", C.c_int),
                ("output_step", C.c_int)]
