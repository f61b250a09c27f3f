"""
Parameter tables and model metadata for the models the hip backend runs.

Restates the parts of the reference's parameter handling that the I(q)
path depends on (sasmodels/modelinfo.py):

* common parameters scale/background (:49-56) and the magnetic parameters
  appended to the call list (:634-672);
* the values-vector layout ``[scale, background, kernel pars...,
  up_frac_i, up_frac_f, up_theta, up_phi, (M0, mtheta, mphi) per sld]``
  and ``nvalues`` (:454-459);
* ``polydisperse`` for volume and orientation parameters and
  ``relative_pd`` for volume parameters (:176-177);
* ``theta_offset``, ``max_pd``, ``pd_1d``/``pd_2d`` (:470-495);
* choice-list limits ``[[name, ...]]`` -> ``(0, len - 1)`` (:95-108);
* vector parameters ``id[ctl]`` (:135-180, :585-606): ``ctl``'s upper limit
  is the length, ``ctl`` becomes a non-polydisperse control parameter, and
  the call list holds one slot per element, ``id1`` .. ``idN`` (:634-647).

The parameter tables are the reference's own (sphere.py:66-72,
cylinder.py:135-142, ellipsoid.py:145-152, core_shell_cylinder.py:118-134,
core_shell_sphere.py:78-82, parallelepiped.py:216-231,
triaxial_ellipsoid.py:143-158, bcc_paracrystal.py:186-194,
fcc_paracrystal.py:184-192, sc_paracrystal.py:181-189, hollow_cylinder.py:90-98,
vesicle.py:87-96, fuzzy_sphere.py:113-117, core_multi_shell.py:91-103,
onion.py:323-333, core_shell_ellipsoid.py:139-149,
elliptical_cylinder.py:115-122, core_shell_bicelle.py:140-151, guinier.py, gaussian_peak.py,
lorentz.py, dab.py, gel_fit.py, mass_fractal.py, surface_fractal.py,
mass_surface_fractal.py, star_polymer.py, mono_gauss_coil.py, fractal.py,
fractal_core_shell.py, lamellar_hg.py, lamellar_stack_caille.py,
lamellar_hg_stack_caille.py, lamellar_stack_paracrystal.py, linear_pearls.py,
polymer_micelle.py, raspberry.py (the ``parameters`` tables), hardsphere.py:84-88,
stickyhardsphere.py:110-121, squarewell.py:83-92, hayter_msa.py:110-117),
with the attributes the
kernels use (``have_Fq``, ``valid``, ``radius_effective_modes``, the
orientation mode, ``structure_factor``).
"""

import numpy as np
from numpy import inf

#: Maximum number of simultaneously polydisperse parameters (modelinfo.py:38)
MAX_PD = 5
NUM_COMMON_PARS = 2
NUM_MAGFIELD_PARS = 4
NUM_MAGNETIC_PARS = 3
DEFAULT_BACKGROUND = 1e-3

COMMON_PARAMETERS = [
    ("scale", "", 1, (0.0, np.inf), "", "Scale factor or Volume fraction"),
    ("background", "1/cm", DEFAULT_BACKGROUND, (-np.inf, np.inf), "", "Source background"),
]


class Parameter:
    """
    One model parameter: *name*, *units*, *default*, *limits*, *type*
    ('volume', 'orientation', 'sld', 'magnetic' or '') and *description*.

    A vector parameter is named ``id[k]`` (fixed length k) or ``id[ctl]``,
    whose length is the upper limit of the control parameter *ctl*; ``id``
    drops the brackets and ``length``/``length_control`` keep the rest
    (modelinfo.py:135-180).
    """
    def __init__(self, name, units="", default=None, limits=(-np.inf, np.inf),
                 type="", description=""):
        self.name = name
        self.id, self.length, self.length_control = _split_vector_name(name)
        self.units = units
        self.default = default
        self.limits = tuple(limits)
        self.type = type
        self.description = description
        self.is_control = False
        self.polydisperse = type in ("volume", "orientation")
        self.relative_pd = type == "volume"
        self.choices = []

    def __repr__(self):
        return "P<%s>" % self.name


def _split_vector_name(name):
    """``(id, length, length_control)`` of a parameter name."""
    if "[" not in name:
        return name, 1, None
    if not name.endswith("]"):
        raise ValueError("Expected name[len] for vector parameter %s" % name)
    pid, ref = name[:-1].split("[", 1)
    ref = ref.strip()
    if not ref:
        raise ValueError("Need to specify vector length for %s" % name)
    if ref.isdigit():
        return pid, int(ref), None
    # the length comes from the control parameter's upper limit once the
    # whole table is known (ParameterTable)
    return pid, 1, ref


def parse_parameter(name, units, default, user_limits, ptype, description):
    """
    Build a :class:`Parameter` from a table row, checking limits and type
    (modelinfo.py:80-181).
    """
    if ptype not in ("volume", "orientation", "sld", "magnetic", ""):
        raise ValueError("unexpected type %r for %s" % (ptype, name))
    choices = []
    if isinstance(user_limits[0], (tuple, list)):
        choices = list(user_limits[0])
        if not all(isinstance(k, str) for k in choices):
            raise ValueError("choices must be strings for %s" % name)
        limits = (0., len(choices) - 1.)
    else:
        low, high = user_limits
        limits = (float(low), float(high))
        if limits[0] >= limits[1]:
            raise ValueError("require lower limit < upper limit")
    if default < limits[0] or default > limits[1]:
        raise ValueError("default value %r not in range for %s" % (default, name))
    pid = _split_vector_name(name)[0]
    if ptype == "" and (pid.startswith("sld") or pid.endswith("sld")):
        ptype = "sld"
    par = Parameter(name, units, float(default), limits, ptype, description)
    par.choices = choices
    return par


def expand_pars(table, pars=None):
    """
    Defaults of *table* updated by *pars*, with a vector parameter given
    as a whole expanded to its slots: ``thickness=[20, 10]`` sets thickness1
    and thickness2, ``thickness=5`` sets every slot not named on its own
    (modelinfo.py:183-225).
    """
    result = dict(table.defaults)
    if not pars:
        return result
    vector = {p.id: p for p in table.kernel_parameters if p.length > 1}
    scalars = {k: v for k, v in pars.items() if k not in vector}
    for name, value in pars.items():
        if name not in vector:
            continue
        if np.isscalar(value):
            for k in range(1, vector[name].length + 1):
                scalars.setdefault(name + str(k), value)
        else:
            for k, v in enumerate(value, 1):
                scalars[name + str(k)] = v
    result.update(scalars)
    return result


class ParameterTable:
    """
    Kernel and call parameter lists for one model (modelinfo.py:390-779).
    """
    def __init__(self, parameters, allow_magnetic=True):
        self.common_parameters = [Parameter(*p) for p in COMMON_PARAMETERS]
        self.kernel_parameters = parameters
        self._set_vector_lengths()
        self.npars = sum(p.length for p in parameters)
        self.nmagnetic = (sum(p.length for p in parameters if p.type == "sld")
                          if allow_magnetic else 0)
        self.nvalues = NUM_COMMON_PARS + self.npars
        if self.nmagnetic:
            self.nvalues += NUM_MAGFIELD_PARS + NUM_MAGNETIC_PARS*self.nmagnetic
        self.call_parameters = self._get_call_parameters()
        self.defaults = {p.id: p.default for p in self.call_parameters}
        self.iq_parameters = [p for p in parameters
                              if p.type not in ("orientation", "magnetic")]
        self.orientation_parameters = [p for p in parameters if p.type == "orientation"]
        self.form_volume_parameters = [p for p in parameters if p.type == "volume"]
        offset = 0
        self.theta_offset = -1
        for p in parameters:
            if p.name == "theta":
                self.theta_offset = offset
                break
            offset += p.length
        num_pd = sum(p.length for p in parameters if p.polydisperse)
        self.max_pd = min(num_pd, MAX_PD)
        self.has_2d = any(p.type in ("orientation", "magnetic") for p in parameters)
        self.is_asymmetric = any(p.name == "psi" for p in parameters)
        self.magnetism_index = [k for k, p in enumerate(self.call_parameters)
                                if p.id.endswith("_M0")]
        self.pd_1d = set(p.name for p in self.call_parameters
                         if p.polydisperse and p.type not in ("orientation", "magnetic"))
        self.pd_2d = set(p.name for p in self.call_parameters if p.polydisperse)

    def _set_vector_lengths(self):
        """
        Length of each ``id[ctl]`` parameter from the upper limit of *ctl*,
        which becomes a control parameter: an integer in [0, 20], never
        polydisperse (modelinfo.py:585-606).
        """
        by_id = {p.id: p for p in self.kernel_parameters}
        for p in self.kernel_parameters:
            if p.length_control is None:
                continue
            ctl = by_id.get(p.length_control)
            if ctl is None:
                raise ValueError("no reference variable %r for %s" % (p.length_control, p.name))
            low, high = ctl.limits
            if int(low) != low or int(high) != high or low < 0 or high > 20:
                raise ValueError("expected limits on %s to be within [0, 20]" % ctl.name)
            ctl.is_control = True
            ctl.polydisperse = False
            p.length = int(high)

    def _get_call_parameters(self):
        # a vector parameter takes one call slot per element, id1 .. idN,
        # with the element's distribution flags (modelinfo.py:634-647)
        full = self.common_parameters[:]
        for p in self.kernel_parameters:
            if p.length == 1:
                full.append(p)
                continue
            for k in range(1, p.length + 1):
                pk = Parameter(p.id + str(k), p.units, p.default, p.limits, p.type, p.description)
                pk.polydisperse, pk.relative_pd, pk.choices = p.polydisperse, p.relative_pd, p.choices
                full.append(pk)
        if self.nmagnetic > 0:
            full.extend([
                Parameter("up_frac_i", "", 0., [0., 1.], "magnetic",
                          "fraction of spin up incident"),
                Parameter("up_frac_f", "", 0., [0., 1.], "magnetic",
                          "fraction of spin up final"),
                Parameter("up_theta", "degrees", 90., [-180., 180.], "magnetic",
                          "polarization axis rotation angle"),
                Parameter("up_phi", "degrees", 0., [-360., 360.], "magnetic",
                          "polarization axis inclination angle"),
            ])
            for p in [p for p in full if p.type == "sld"]:
                full.extend([
                    Parameter(p.id + "_M0", "1e-6/Ang^2", 0., [-np.inf, np.inf],
                              "magnetic", "magnetic amplitude for " + p.description),
                    Parameter(p.id + "_mtheta", "degrees", 0., [-180., 180.],
                              "magnetic", "magnetic latitude for " + p.description),
                    Parameter(p.id + "_mphi", "degrees", 0., [-360., 360.],
                              "magnetic", "magnetic longitude for " + p.description),
                ])
        return full

    def __getitem__(self, key):
        for par in self.call_parameters:
            if par.name == key:
                return par
        raise KeyError("unknown parameter %r" % key)

    def __contains__(self, key):
        return any(par.name == key for par in self.call_parameters)


def make_parameter_table(pars, allow_magnetic=True):
    """Build a :class:`ParameterTable` from 6-element table rows."""
    processed = []
    for p in pars:
        if not isinstance(p, (list, tuple)) or len(p) != 6:
            raise ValueError("Parameter should be [name, units, default, limits, type, desc], but got %r" % (p,))
        processed.append(parse_parameter(*p))
    return ParameterTable(processed, allow_magnetic=allow_magnetic)


class ModelInfo:
    """
    Model metadata used by the kernels (the fields of modelinfo.py:1033-1228
    that the I(q) path reads).
    """
    id = None
    name = None
    title = None
    parameters = None
    base = None
    composition = None
    compiled = True
    have_Fq = False
    valid = ""
    radius_effective_modes = None
    structure_factor = False
    xy_mode = "qa"
    #: id of the model in the C ABI (include/saship.h)
    kernel_id = None
    #: Gauss-Legendre size for the orientation average; 0 is the reference's
    #: adaptive choice (models/lib/adaptive.c:11276-11298)
    integration_size = 0
    #: models whose 1D kernel integrates over orientation
    has_quadrature = False
    #: False for models the reference runs in double precision only
    single = True

    def get_hidden_parameters(self, control):
        """
        Call parameters a model with one control parameter ignores at
        control value *control*: the vector slots above it and their
        magnetic parameters (modelinfo.py:1205-1228).  Empty without vectors.
        """
        table = self.parameters
        vectors = [p for p in table.kernel_parameters if p.length > 1]
        if not vectors:
            return set()
        if sum(p.is_control for p in table.kernel_parameters) != 1:
            raise ValueError("more than one control parameter")
        hidden = set()
        for p in vectors:
            for k in range(int(control) + 1, p.length + 1):
                hidden.add(p.id + str(k))
                if p.type == "sld":
                    hidden.update(p.id + str(k) + s for s in ("_M0", "_mtheta", "_mphi"))
        return hidden


def _paracrystal(kernel_id, dnn_limits, sld, sld_solvent, angle, names):
    return (kernel_id, [
        ["dnn", "Ang", 220.0, dnn_limits, "", "Nearest neighbour distance"],
        ["d_factor", "", 0.06, [-inf, inf], "", "Paracrystal distortion factor"],
        ["radius", "Ang", 40.0, [0, inf], "volume", names[0]],
        ["sld", "1e-6/Ang^2", sld[0], sld[1], "sld", names[1]],
        ["sld_solvent", "1e-6/Ang^2", sld_solvent[0], sld_solvent[1], "sld", "Solvent scattering length density"],
        ["theta", "degrees", angle, [-180, 180], "orientation", "c axis to beam angle"],
        ["phi", "degrees", angle, [-360, 360], "orientation", "rotation about beam"],
        ["psi", "degrees", angle, [-360, 360], "orientation", "rotation about c axis"],
    ], dict(have_Fq=False, xy_mode="qabc", has_quadrature=True))


# name -> (kernel id, parameter rows, extra attributes)
_MODEL_TABLE = {
    "sphere": (0, [
        ["sld", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Layer scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6, [-inf, inf], "sld", "Solvent scattering length density"],
        ["radius", "Ang", 50, [0, inf], "volume", "Sphere radius"],
    ], dict(radius_effective_modes=["radius"], xy_mode="qa", has_quadrature=False)),
    "cylinder": (1, [
        ["sld", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Cylinder scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent scattering length density"],
        ["radius", "Ang", 20, [0, inf], "volume", "Cylinder radius"],
        ["length", "Ang", 400, [0, inf], "volume", "Cylinder length"],
        ["theta", "degrees", 60, [-180, 180], "orientation", "cylinder axis to beam angle"],
        ["phi", "degrees", 60, [-360, 360], "orientation", "rotation about beam"],
    ], dict(valid="radius >= 0.0 && length >= 0.0",
            radius_effective_modes=[
                "excluded volume", "equivalent volume sphere", "radius",
                "half length", "half min dimension", "half max dimension",
                "half diagonal"],
            xy_mode="qac", has_quadrature=True)),
    "ellipsoid": (2, [
        ["sld", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Ellipsoid scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent scattering length density"],
        ["radius_polar", "Ang", 20, [0, inf], "volume", "Polar radius"],
        ["radius_equatorial", "Ang", 400, [0, inf], "volume", "Equatorial radius"],
        ["theta", "degrees", 60, [-180, 180], "orientation", "ellipsoid axis to beam angle"],
        ["phi", "degrees", 60, [-360, 360], "orientation", "rotation about beam"],
    ], dict(radius_effective_modes=[
                "average curvature", "equivalent volume sphere", "min radius",
                "max radius"],
            xy_mode="qac", has_quadrature=True)),
    "core_shell_cylinder": (3, [
        ["sld_core", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Cylinder core scattering length density"],
        ["sld_shell", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Cylinder shell scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent scattering length density"],
        ["radius", "Ang", 20, [0, inf], "volume", "Cylinder core radius"],
        ["thickness", "Ang", 20, [0, inf], "volume", "Cylinder shell thickness"],
        ["length", "Ang", 400, [0, inf], "volume", "Cylinder length"],
        ["theta", "degrees", 60, [-180, 180], "orientation", "cylinder axis to beam angle"],
        ["phi", "degrees", 60, [-360, 360], "orientation", "rotation about beam"],
    ], dict(radius_effective_modes=[
                "excluded volume", "equivalent volume sphere", "outer radius",
                "half outer length", "half min outer dimension",
                "half max outer dimension", "half outer diagonal"],
            xy_mode="qac", has_quadrature=True)),
    "core_shell_sphere": (4, [
        ["radius", "Ang", 60.0, [0, inf], "volume", "Sphere core radius"],
        ["thickness", "Ang", 10.0, [0, inf], "volume", "Sphere shell thickness"],
        ["sld_core", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "core scattering length density"],
        ["sld_shell", "1e-6/Ang^2", 2.0, [-inf, inf], "sld", "shell scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 3.0, [-inf, inf], "sld", "Solvent scattering length density"],
    ], dict(radius_effective_modes=["outer radius", "core radius"], xy_mode="qa",
            has_quadrature=False)),
    "parallelepiped": (6, [
        ["sld", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Parallelepiped scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent scattering length density"],
        ["length_a", "Ang", 35, [0, inf], "volume", "Shorter side of the parallelepiped"],
        ["length_b", "Ang", 75, [0, inf], "volume", "Second side of the parallelepiped"],
        ["length_c", "Ang", 400, [0, inf], "volume", "Larger side of the parallelepiped"],
        ["theta", "degrees", 60, [-180, 180], "orientation", "c axis to beam angle"],
        ["phi", "degrees", 60, [-360, 360], "orientation", "rotation about beam"],
        ["psi", "degrees", 60, [-360, 360], "orientation", "rotation about c axis"],
    ], dict(radius_effective_modes=[
                "equivalent cylinder excluded volume", "equivalent volume sphere",
                "half length_a", "half length_b", "half length_c",
                "equivalent circular cross-section", "half ab diagonal", "half diagonal"],
            xy_mode="qabc", has_quadrature=True)),
    "triaxial_ellipsoid": (7, [
        ["sld", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Ellipsoid scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent scattering length density"],
        ["radius_equat_minor", "Ang", 20, [0, inf], "volume", "Minor equatorial radius, Ra"],
        ["radius_equat_major", "Ang", 400, [0, inf], "volume", "Major equatorial radius, Rb"],
        ["radius_polar", "Ang", 10, [0, inf], "volume", "Polar radius, Rc"],
        ["theta", "degrees", 60, [-180, 180], "orientation", "polar axis to beam angle"],
        ["phi", "degrees", 60, [-360, 360], "orientation", "rotation about beam"],
        ["psi", "degrees", 60, [-360, 360], "orientation", "rotation about polar axis"],
    ], dict(radius_effective_modes=[
                "equivalent biaxial ellipsoid average curvature", "equivalent volume sphere",
                "min radius", "max radius"],
            xy_mode="qabc", has_quadrature=True)),
    # spheres on distorted lattices: Iq only, FP64, Gauss-150 product rule
    # over (theta, phi) unless an integration size is set
    "bcc_paracrystal": _paracrystal(8, [-inf, inf], (4, [-inf, inf]), (1, [-inf, inf]), 60,
                                    ("Particle radius", "Particle scattering length density")),
    "fcc_paracrystal": _paracrystal(9, [-inf, inf], (4, [-inf, inf]), (1, [-inf, inf]), 60,
                                    ("Particle radius", "Particle scattering length density")),
    "sc_paracrystal": _paracrystal(10, [0.0, inf], (3.0, [0.0, inf]), (6.3, [0.0, inf]), 0,
                                   ("Radius of sphere", "Sphere scattering length density")),
    "hollow_cylinder": (11, [
        ["radius", "Ang", 20.0, [0, inf], "volume", "Cylinder core radius"],
        ["thickness", "Ang", 10.0, [0, inf], "volume", "Cylinder wall thickness"],
        ["length", "Ang", 400.0, [0, inf], "volume", "Cylinder total length"],
        ["sld", "1e-6/Ang^2", 6.3, [-inf, inf], "sld", "Cylinder sld"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent sld"],
        ["theta", "degrees", 90, [-180, 180], "orientation", "Cylinder axis to beam angle"],
        ["phi", "degrees", 0, [-360, 360], "orientation", "Rotation about beam"],
    ], dict(radius_effective_modes=[
                "excluded volume", "equivalent outer volume sphere", "outer radius", "half length",
                "half outer min dimension", "half outer max dimension", "half outer diagonal"],
            xy_mode="qac", has_quadrature=True)),
    "vesicle": (12, [
        ["sld", "1e-6/Ang^2", 0.5, [-inf, inf], "sld", "vesicle shell scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6.36, [-inf, inf], "sld", "solvent scattering length density"],
        ["volfraction", "", 0.05, [0, 1.0], "", "volume fraction of vesicles"],
        ["radius", "Ang", 100, [0, inf], "volume", "radius of solvent filled core"],
        ["thickness", "Ang", 30, [0, inf], "volume", "thickness of vesicle shell"],
    ], dict(radius_effective_modes=["outer radius"], xy_mode="qa", has_quadrature=False)),
    "fuzzy_sphere": (13, [
        ["sld", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Scattering length density of sphere"],
        ["sld_solvent", "1e-6/Ang^2", 3, [-inf, inf], "sld", "Scattering length density of solvent"],
        ["radius", "Ang", 60, [0, inf], "volume", "Pseudo-radius of sphere"],
        ["fuzziness", "Ang", 10, [0, inf], "volume", "Std dev of the fuzzy interface"],
    ], dict(radius_effective_modes=["radius", "radius + fuzziness"], xy_mode="qa",
            has_quadrature=False)),
    # concentric shells: n (a control parameter, integer in [0, 10]) sets how
    # many of the ten sld/thickness slots are used
    "core_multi_shell": (17, [
        ["sld_core", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "Core scattering length density"],
        ["radius", "Ang", 200., [0, inf], "volume", "Radius of the core"],
        ["sld_solvent", "1e-6/Ang^2", 6.4, [-inf, inf], "sld", "Solvent scattering length density"],
        ["n", "", 1, [0, 10], "volume", "number of shells"],
        ["sld[n]", "1e-6/Ang^2", 1.7, [-inf, inf], "sld", "scattering length density of shell k"],
        ["thickness[n]", "Ang", 40., [0, inf], "volume", "Thickness of shell k"],
    ], dict(radius_effective_modes=["outer radius", "core radius"], xy_mode="qa", has_quadrature=False)),
    # shells with exponential sld profiles; FP64 like the reference (single = False)
    "onion": (18, [
        ["sld_core", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "Core scattering length density"],
        ["radius_core", "Ang", 200., [0, inf], "volume", "Radius of the core"],
        ["sld_solvent", "1e-6/Ang^2", 6.4, [-inf, inf], "sld", "Solvent scattering length density"],
        ["n_shells", "", 1, [0, 10], "volume", "number of shells (must be integer)"],
        ["sld_in[n_shells]", "1e-6/Ang^2", 1.7, [-inf, inf], "sld",
         "scattering length density at the inner radius of shell k"],
        ["sld_out[n_shells]", "1e-6/Ang^2", 2.0, [-inf, inf], "sld",
         "scattering length density at the outer radius of shell k"],
        ["thickness[n_shells]", "Ang", 40., [0, inf], "volume", "Thickness of shell k"],
        ["A[n_shells]", "", 1.0, [-inf, inf], "", "Decay rate of shell k"],
    ], dict(radius_effective_modes=["outer radius"], xy_mode="qa", has_quadrature=False, single=False)),
    "core_shell_ellipsoid": (19, [
        ["radius_equat_core", "Ang", 20, [0, inf], "volume", "Equatorial radius of core"],
        ["x_core", "None", 3, [0, inf], "volume", "axial ratio of core, X = r_polar/r_equatorial"],
        ["thick_shell", "Ang", 30, [0, inf], "volume", "thickness of shell at equator"],
        ["x_polar_shell", "", 1, [0, inf], "volume", "ratio of thickness of shell at pole to that at equator"],
        ["sld_core", "1e-6/Ang^2", 2, [-inf, inf], "sld", "Core scattering length density"],
        ["sld_shell", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Shell scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6.3, [-inf, inf], "sld", "Solvent scattering length density"],
        ["theta", "degrees", 0, [-180, 180], "orientation", "elipsoid axis to beam angle"],
        ["phi", "degrees", 0, [-360, 360], "orientation", "rotation about beam"],
    ], dict(radius_effective_modes=["average outer curvature", "equivalent volume sphere",
                                    "min outer radius", "max outer radius"],
            xy_mode="qac", has_quadrature=True)),
    "elliptical_cylinder": (20, [
        ["radius_minor", "Ang", 20.0, [0, inf], "volume", "Ellipse minor radius"],
        ["axis_ratio", "", 1.5, [1, inf], "volume", "Ratio of major radius over minor radius"],
        ["length", "Ang", 400.0, [1, inf], "volume", "Length of the cylinder"],
        ["sld", "1e-6/Ang^2", 4.0, [-inf, inf], "sld", "Cylinder scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "Solvent scattering length density"],
        ["theta", "degrees", 90.0, [-180, 180], "orientation", "cylinder axis to beam angle"],
        ["phi", "degrees", 0, [-360, 360], "orientation", "rotation about beam"],
        ["psi", "degrees", 0, [-360, 360], "orientation", "rotation about cylinder axis"],
    ], dict(radius_effective_modes=[
                "equivalent cylinder excluded volume", "equivalent volume sphere", "average radius",
                "min radius", "max radius", "equivalent circular cross-section", "half length",
                "half min dimension", "half max dimension", "half diagonal"],
            xy_mode="qabc", has_quadrature=True)),
    "core_shell_bicelle": (21, [
        ["radius", "Ang", 80, [0, inf], "volume", "Cylinder core radius"],
        ["thick_rim", "Ang", 10, [0, inf], "volume", "Rim shell thickness"],
        ["thick_face", "Ang", 10, [0, inf], "volume", "Cylinder face thickness"],
        ["length", "Ang", 50, [0, inf], "volume", "Cylinder length"],
        ["sld_core", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Cylinder core scattering length density"],
        ["sld_face", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Cylinder face scattering length density"],
        ["sld_rim", "1e-6/Ang^2", 4, [-inf, inf], "sld", "Cylinder rim scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1, [-inf, inf], "sld", "Solvent scattering length density"],
        ["theta", "degrees", 90, [-180, 180], "orientation", "cylinder axis to beam angle"],
        ["phi", "degrees", 0, [-360, 360], "orientation", "rotation about beam"],
    ], dict(radius_effective_modes=["excluded volume", "equivalent volume sphere", "outer rim radius",
                                    "half outer thickness", "half diagonal"],
            xy_mode="qac", has_quadrature=True)),
    # models defined by I(q) alone (no <F>), evaluated in FP64
    "guinier": (22, [
        ["rg", "Ang", 60.0, [-inf, inf], "", "Radius of Gyration"],
    ], dict(have_Fq=False)),
    "gaussian_peak": (23, [
        ["peak_pos", "1/Ang", 0.05, [-inf, inf], "", "Peak position"],
        ["sigma", "1/Ang", 0.005, [0, inf], "", "Peak width (standard deviation)"],
    ], dict(have_Fq=False)),
    "lorentz": (24, [
        ["cor_length", "Ang", 50.0, [0, inf], "", "Screening length"],
    ], dict(have_Fq=False)),
    "dab": (25, [
        ["cor_length", "Ang", 50.0, [0, inf], "", "correlation length"],
    ], dict(have_Fq=False)),
    "gel_fit": (26, [
        ["guinier_scale", "cm^-1", 1.7, [-inf, inf], "", "Guinier term scale"],
        ["lorentz_scale", "cm^-1", 3.5, [-inf, inf], "", "Lorentz term scale"],
        ["rg", "Ang", 104.0, [2, inf], "", "Radius of gyration"],
        ["fractal_dim", "", 2.0, [0, inf], "", "Fractal exponent"],
        ["cor_length", "Ang", 16.0, [0, inf], "", "Correlation length"],
    ], dict(have_Fq=False)),
    "mass_fractal": (27, [
        ["radius", "Ang", 10.0, [0.0, inf], "", "Particle radius"],
        ["fractal_dim_mass", "", 1.9, [1.0, 6.0], "", "Mass fractal dimension"],
        ["cutoff_length", "Ang", 100.0, [0.0, inf], "", "Cut-off length"],
    ], dict(have_Fq=False)),
    "surface_fractal": (28, [
        ["radius", "Ang", 10.0, [0, inf], "", "Particle radius"],
        ["fractal_dim_surf", "", 2.0, [1, 3], "", "Surface fractal dimension"],
        ["cutoff_length", "Ang", 500., [0.0, inf], "", "Cut-off Length"],
    ], dict(have_Fq=False)),
    "mass_surface_fractal": (29, [
        ["fractal_dim_mass", "", 1.8, [0.0, 6.0], "", "Mass fractal dimension"],
        ["fractal_dim_surf", "", 2.3, [0.0, 6.0], "", "Surface fractal dimension"],
        ["rg_cluster", "Ang", 4000., [0.0, inf], "", "Cluster radius of gyration"],
        ["rg_primary", "Ang", 86.7, [0.0, inf], "", "Primary particle radius of gyration"],
    ], dict(have_Fq=False)),
    "star_polymer": (30, [
        ["rg_squared", "Ang^2", 100.0, [0.0, inf], "", "Ensemble radius of gyration SQUARED of the full polymer"],
        ["arms", "", 3, [1.0, 6.0], "", "Number of arms in the model (integer)"],
    ], dict(have_Fq=False)),
    "mono_gauss_coil": (31, [
        ["i_zero", "1/cm", 70.0, [0.0, inf], "", "Intensity at q=0"],
        ["rg", "Ang", 75.0, [0.0, inf], "volume", "Radius of gyration"],
    ], dict(have_Fq=False, radius_effective_modes=["R_g", "2R_g", "3R_g", "sqrt(5/3)*R_g"])),
    "fractal": (32, [
        ["volfraction", "", 0.05, [0.0, 1], "", "volume fraction of blocks"],
        ["radius", "Ang", 5.0, [0.0, inf], "volume", "radius of particles"],
        ["fractal_dim", "", 2.0, [0.0, 6.0], "", "fractal dimension"],
        ["cor_length", "Ang", 100.0, [0.0, inf], "", "cluster correlation length"],
        ["sld_block", "1e-6/Ang^2", 2.0, [-inf, inf], "sld", "scattering length density of particles"],
        ["sld_solvent", "1e-6/Ang^2", 6.4, [-inf, inf], "sld", "scattering length density of solvent"],
    ], dict(have_Fq=False)),
    "fractal_core_shell": (33, [
        ["radius", "Ang", 60.0, [0.0, inf], "volume", "Sphere core radius"],
        ["thickness", "Ang", 10.0, [0.0, inf], "volume", "Sphere shell thickness"],
        ["sld_core", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "Sphere core scattering length density"],
        ["sld_shell", "1e-6/Ang^2", 2.0, [-inf, inf], "sld", "Sphere shell scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 3.0, [-inf, inf], "sld", "Solvent scattering length density"],
        ["volfraction", "", 0.05, [0.0, inf], "", "Volume fraction of building block spheres"],
        ["fractal_dim", "", 2.0, [0.0, 6.0], "", "Fractal dimension"],
        ["cor_length", "Ang", 100.0, [0.0, inf], "", "Correlation length of fractal-like aggregates"],
    ], dict(have_Fq=False)),
    # the lamellar models have volume parameters but a form volume of 1
    "lamellar_hg": (34, [
        ["length_tail", "Ang", 15, [0, inf], "volume", "Tail thickness ( total = H+T+T+H)"],
        ["length_head", "Ang", 10, [0, inf], "volume", "Head thickness"],
        ["sld", "1e-6/Ang^2", 0.4, [-inf, inf], "sld", "Tail scattering length density"],
        ["sld_head", "1e-6/Ang^2", 3.0, [-inf, inf], "sld", "Head scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6, [-inf, inf], "sld", "Solvent scattering length density"],
    ], dict(have_Fq=False)),
    "lamellar_stack_caille": (35, [
        ["thickness", "Ang", 30.0, [0, inf], "volume", "sheet thickness"],
        ["Nlayers", "", 20, [1, inf], "", "Number of layers"],
        ["d_spacing", "Ang", 400., [0.0, inf], "volume", "lamellar d-spacing of Caille S(Q)"],
        ["Caille_parameter", "1/Ang^2", 0.1, [0.0, 0.8], "", "Caille parameter"],
        ["sld", "1e-6/Ang^2", 6.3, [-inf, inf], "sld", "layer scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "Solvent scattering length density"],
    ], dict(have_Fq=False, single=False)),
    "lamellar_hg_stack_caille": (36, [
        ["length_tail", "Ang", 10, [0, inf], "volume", "Tail thickness"],
        ["length_head", "Ang", 2, [0, inf], "volume", "head thickness"],
        ["Nlayers", "", 30, [1, inf], "", "Number of layers"],
        ["d_spacing", "Ang", 40., [0.0, inf], "volume", "lamellar d-spacing of Caille S(Q)"],
        ["Caille_parameter", "", 0.001, [0.0, 0.8], "", "Caille parameter"],
        ["sld", "1e-6/Ang^2", 0.4, [-inf, inf], "sld", "Tail scattering length density"],
        ["sld_head", "1e-6/Ang^2", 2.0, [-inf, inf], "sld", "Head scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6, [-inf, inf], "sld", "Solvent scattering length density"],
    ], dict(have_Fq=False, single=False)),
    "lamellar_stack_paracrystal": (37, [
        ["thickness", "Ang", 33.0, [0, inf], "volume", "sheet thickness"],
        ["Nlayers", "", 20, [1, inf], "", "Number of layers"],
        ["d_spacing", "Ang", 250., [0.0, inf], "", "lamellar spacing of paracrystal stack"],
        ["sigma_d", "Ang", 0.0, [0.0, inf], "", "Sigma (polydispersity) of the lamellar spacing"],
        ["sld", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "layer scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6.34, [-inf, inf], "sld", "Solvent scattering length density"],
    ], dict(have_Fq=False, single=False)),
    "linear_pearls": (38, [
        ["radius", "Ang", 80.0, [0, inf], "", "Radius of the pearls"],
        ["edge_sep", "Ang", 350.0, [0, inf], "", "Length of the string segment - surface to surface"],
        ["num_pearls", "", 3.0, [1, inf], "", "Number of the pearls"],
        ["sld", "1e-6/Ang^2", 1.0, [-inf, inf], "sld", "SLD of the pearl spheres"],
        ["sld_solvent", "1e-6/Ang^2", 6.3, [-inf, inf], "sld", "SLD of the solvent"],
    ], dict(have_Fq=False, single=False)),
    "polymer_micelle": (39, [
        ["ndensity", "1e15/cm^3", 8.94, [0.0, inf], "", "Number density of micelles"],
        ["v_core", "Ang^3", 62624.0, [0.0, inf], "", "Core volume "],
        ["v_corona", "Ang^3", 61940.0, [0.0, inf], "", "Corona volume"],
        ["sld_solvent", "1e-6/Ang^2", 6.4, [0.0, inf], "sld", "Solvent scattering length density"],
        ["sld_core", "1e-6/Ang^2", 0.34, [0.0, inf], "sld", "Core scattering length density"],
        ["sld_corona", "1e-6/Ang^2", 0.8, [0.0, inf], "sld", "Corona scattering length density"],
        ["radius_core", "Ang", 45.0, [0.0, inf], "", "Radius of core ( must be >> rg )"],
        ["rg", "Ang", 20.0, [0.0, inf], "", "Radius of gyration of chains in corona"],
        ["d_penetration", "", 1.0, [-inf, inf], "", "Factor to mimic non-penetration of Gaussian chains"],
        ["n_aggreg", "", 6.0, [-inf, inf], "", "Aggregation number of the micelle"],
    ], dict(have_Fq=False, single=False)),
    "raspberry": (40, [
        ["sld_lg", "1e-6/Ang^2", -0.4, [-inf, inf], "sld", "large particle scattering length density"],
        ["sld_sm", "1e-6/Ang^2", 3.5, [-inf, inf], "sld", "small particle scattering length density"],
        ["sld_solvent", "1e-6/Ang^2", 6.36, [-inf, inf], "sld", "solvent scattering length density"],
        ["volfraction_lg", "", 0.05, [-inf, inf], "", "volume fraction of large spheres"],
        ["volfraction_sm", "", 0.005, [-inf, inf], "", "volume fraction of small spheres"],
        ["surface_fraction", "", 0.4, [-inf, inf], "", "fraction of small spheres at surface"],
        ["radius_lg", "Ang", 5000, [0, inf], "volume", "radius of large spheres"],
        ["radius_sm", "Ang", 100, [0, inf], "volume", "radius of small spheres"],
        ["penetration", "Ang", 0, [-1, 1], "volume", "fractional penetration depth of small spheres into large sphere"],
    ], dict(have_Fq=False, radius_effective_modes=["radius_large", "radius_outer"])),
    # structure factor: Iq only (no F), no volume, nothing polydisperse;
    # evaluated in FP64 (the reference marks it double precision only)
    "hardsphere": (5, [
        ["radius_effective", "Ang", 50.0, [0, inf], "", "effective radius of hard sphere"],
        ["volfraction", "", 0.2, [0, 0.74], "", "volume fraction of hard spheres"],
    ], dict(structure_factor=True, have_Fq=False, xy_mode="qa", has_quadrature=False)),
    "stickyhardsphere": (14, [
        ["radius_effective", "Ang", 50.0, [0, inf], "", "effective radius of hard sphere"],
        ["volfraction", "", 0.2, [0, 0.74], "", "volume fraction of hard spheres"],
        ["perturb", "", 0.05, [0.01, 0.1], "", "perturbation parameter, tau"],
        ["stickiness", "", 0.20, [-inf, inf], "", "stickiness, epsilon"],
    ], dict(structure_factor=True, have_Fq=False, xy_mode="qa", has_quadrature=False)),
    "squarewell": (15, [
        ["radius_effective", "Ang", 50.0, [0, inf], "", "effective radius of hard sphere"],
        ["volfraction", "", 0.04, [0, 0.08], "", "volume fraction of spheres"],
        ["welldepth", "kT", 1.5, [0.0, 1.5], "", "depth of well, epsilon"],
        ["wellwidth", "diameters", 1.2, [1.0, inf], "", "width of well in diameters (=2R) units, must be > 1"],
    ], dict(structure_factor=True, have_Fq=False, xy_mode="qa", has_quadrature=False)),
    "hayter_msa": (16, [
        ["radius_effective", "Ang", 20.75, [0, inf], "", "effective radius of charged sphere"],
        ["volfraction", "None", 0.0192, [0, 0.74], "", "volume fraction of spheres"],
        ["charge", "e", 19.0, [0.000001, 200], "", "charge on sphere (in electrons)"],
        ["temperature", "K", 318.16, [0, 450], "", "temperature, in Kelvin, for Debye length calculation"],
        ["concentration_salt", "M", 0.0, [0, inf], "", "conc of salt, moles/litre, 1:1 electolyte, for Debye length"],
        ["dielectconst", "None", 71.08, [-inf, inf], "",
         "dielectric constant (relative permittivity) of solvent, kappa, default water, for Debye length"],
    ], dict(structure_factor=True, have_Fq=False, xy_mode="qa", has_quadrature=False)),
}

MODELS = tuple(_MODEL_TABLE)


def make_model_info(name):
    """
    Return a fresh :class:`ModelInfo` for model *name* (a new object each
    call, so :func:`generate.set_integration_size` on it affects only it).
    """
    if name not in _MODEL_TABLE:
        raise ValueError("model %r is not available on the hip platform; "
                         "available models: %s" % (name, ", ".join(MODELS)))
    kernel_id, rows, extra = _MODEL_TABLE[name]
    info = ModelInfo()
    info.id = name
    info.name = " ".join(w.capitalize() for w in name.split("_"))
    info.title = info.name + " model"
    info.kernel_id = kernel_id
    info.parameters = info.base = make_parameter_table(rows)
    info.have_Fq = True
    info.structure_factor = False
    info.integration_size = 0
    for key, value in extra.items():
        setattr(info, key, value)
    return info
