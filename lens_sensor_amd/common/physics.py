"""Physical constants, material dispersion and the complex-index loss model.

Behaviour follows common/physics.py of the reference:
  constants            common/physics.py:10-13
  Sellmeier models     common/physics.py:43-92
  n_at                 common/physics.py:97-104
  layered cornea       common/physics.py:120-163
  complex index/sigma  common/physics.py:166-220
  core-only source     common/physics.py:226-244
"""
import numpy as np

C0 = 2.99792458e8             # vacuum speed of light [m/s]
EPS0 = 8.8541878128e-12       # vacuum permittivity [F/m]
MU0 = 4.0 * np.pi * 1e-7      # vacuum permeability [H/m]
N_AIR = 1.000

LAM_VIS_UM = 0.5893           # sodium D line: wavelength of the visible literature values
MUCIN_DN = 0.0018             # mucus above tear fluid: 1 % mucin x dn/dc 0.18 mL/g


def _sellmeier(lam_um, coeffs):
    lsq = lam_um * lam_um
    terms = 0
    for b, c in coeffs:
        terms = terms + b * lsq / (lsq - c)
    return float(np.sqrt(1.0 + terms))


# Sellmeier (B, C[um^2]) pairs: PMMA and polystyrene (Sultanova et al. 2009),
# water (Daimon & Masumura 2007).
_SELL = {
    'pmma': ((1.1819, 0.011313),),
    'polystyrol': ((1.4435, 0.020216),),
    'water': ((5.684027565e-1, 5.101829712e-3), (1.726177391e-1, 1.821153936e-2),
              (2.086189578e-2, 2.620722293e-2), (1.130748688e-1, 1.069792721e1)),
}


def _n_water(lam_um):
    return _sellmeier(lam_um, _SELL['water'])


def _n_pmma(lam_um):
    return _sellmeier(lam_um, _SELL['pmma'])


def _offset_model(n_visible, base):
    """n(lambda) = n_visible + base(lambda) - base(visible): keeps the base dispersion."""
    return lambda lam_um: n_visible + base(lam_um) - base(LAM_VIS_UM)


MATERIAL_MODELS = {
    'pmma': _n_pmma,
    'polystyrol': lambda l: _sellmeier(l, _SELL['polystyrol']),
    'water': _n_water,
    'aqueous': lambda l: _n_water(l) + 0.0035,
    'cornea': _offset_model(1.376, _n_water),
    'mucin': lambda l: _n_water(l) + 0.0035 + MUCIN_DN,
    'lipid': _offset_model(1.480, _n_pmma),
    'epithelium': _offset_model(1.401, _n_water),
    'stroma': _offset_model(1.3765, _n_water),
    'aqueous_humor': _offset_model(1.336, _n_water),
}

_FIT_RANGE_UM = (0.437, 1.052)
_noted = set()


def n_at(mat, lam_nm):
    """Refractive index of material ``mat`` at ``lam_nm`` nanometres."""
    lam_um = float(lam_nm) / 1000.0
    if not (_FIT_RANGE_UM[0] <= lam_um <= _FIT_RANGE_UM[1]) and (mat, lam_nm) not in _noted:
        _noted.add((mat, lam_nm))
        print(f'[physics] Note: {mat} at {lam_nm:g} nm is outside the fit range (extrapolated).')
    return float(MATERIAL_MODELS[mat](lam_um))


# Layered central cornea, depth below the mucin/cornea boundary in um (visible values).
CORNEA_LAYERS_UM = dict(epithelium=50.0, bowman=12.0, stroma_end=525.0, endothelium_end=545.0)
_N_EPI, _N_STROMA_FRONT, _N_STROMA_BACK, _N_HUMOR = 1.401, 1.380, 1.373, 1.336


class CorneaProfile(float):
    """Cornea index as a function of depth.  The float value is the epithelium (the
    layer directly below the mucin); material builders call ``at_depth``."""

    def __new__(cls, lam_nm):
        obj = float.__new__(cls, n_at('epithelium', lam_nm))
        obj.lam_nm = float(lam_nm)
        return obj

    def __reduce__(self):
        return (CorneaProfile, (self.lam_nm,))

    def at_depth(self, d_um):
        d = np.asarray(d_um, dtype=np.float64)
        lay = CORNEA_LAYERS_UM
        b_epi = lay['epithelium']
        b_bow = b_epi + lay['bowman']
        n = np.full(d.shape, _N_HUMOR)
        n = np.where(d < lay['endothelium_end'], _N_STROMA_BACK, n)
        in_stroma = (d >= b_bow) & (d < lay['stroma_end'])
        frac = np.clip((d - b_bow) / (lay['stroma_end'] - b_bow), 0.0, 1.0)
        n = np.where(in_stroma, _N_STROMA_FRONT + (_N_STROMA_BACK - _N_STROMA_FRONT) * frac, n)
        n = np.where((d >= b_epi) & (d < b_bow), _N_STROMA_FRONT, n)
        n = np.where(d < b_epi, _N_EPI, n)
        return n + _n_water(self.lam_nm / 1000.0) - _n_water(LAM_VIS_UM)


def cornea_default(lam_nm):
    return CorneaProfile(lam_nm)


def parse_n(s):
    """'1.52' -> 1.52; '1.33+0.001i' / '1.33+1e-3j' -> complex (imag = extinction kappa)."""
    if isinstance(s, (int, float, complex)):
        v = s
    else:
        t = str(s).strip().lower().replace(' ', '').replace('i', 'j')
        if t.startswith('n='):
            t = t[2:]
        v = complex(t) if 'j' in t else float(t)
    if isinstance(v, complex):
        if v.imag < 0:
            raise ValueError(f"kappa < 0 ({s}) would mean gain; convention n = n' + i*kappa")
        if v.imag >= v.real:
            raise ValueError(f"kappa >= n' ({s}): metallic, needs a Drude model (not supported)")
        if v.imag == 0:
            return float(v.real)
    return v


def split_n(v):
    if v is None:
        return None, 0.0
    if isinstance(v, complex):
        return float(v.real), float(v.imag)
    return v, 0.0


def loss_sigma(eps_imag, lam_m):
    """Conductivity [S/m] for Im(eps_r) at the operating wavelength: omega0 eps0 eps''."""
    return 2 * np.pi * C0 / lam_m * EPS0 * np.asarray(eps_imag)


SRC_CORE_RAMP_UM = 0.2


def core_source_mask(y_um, t_wg_um, ramp_um=SRC_CORE_RAMP_UM, z_um=None, wg_width_um=None):
    """Source weight 1 inside the core, cos^2 ramp over ``ramp_um`` outside, else 0."""
    def ramp(d):
        d = np.asarray(d, dtype=np.float64)
        return np.where(d <= 0.0, 1.0,
                        np.where(d < ramp_um, np.cos(0.5 * np.pi * d / max(ramp_um, 1e-9)) ** 2, 0.0))
    y = np.asarray(y_um, dtype=np.float64)
    m = ramp(np.maximum(-y, y - float(t_wg_um)))
    if z_um is not None and wg_width_um:
        m = m * ramp(np.abs(np.asarray(z_um, dtype=np.float64)) - 0.5 * float(wg_width_um))
    return m
