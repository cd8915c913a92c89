"""Layer stack below the waveguide core (planar geometry).

A stack is a list of layers from the core downwards, each
{'name', 't_um', 'n' (complex), 'kappa', 'mat'}; behaviour of common/layer_stack.py of
the reference (parse_layer_stack :47-74, legacy_stack :77-87, stack_total_um :90-91,
stack_bounds :94-100).
"""
import numpy as np

LEGACY_NAMES = ('Lipid', 'Aqueous', 'Mucin')


def _parse_n(tok, lam_nm):
    tok = tok.strip()
    try:
        v = complex(tok.replace(' ', ''))
        return complex(v.real, abs(v.imag)), None
    except ValueError:
        from lens_sensor_amd.common.physics import n_at
        return complex(n_at(tok.lower(), lam_nm), 0.0), tok.lower()


def parse_layer_stack(spec, lam_nm):
    """``Name:thickness_um:n[,...]`` -> stack; 'none' -> []; None -> None."""
    if spec is None:
        return None
    s = str(spec).strip()
    if s.lower() in ('', 'none', '-'):
        return []
    out = []
    for i, part in enumerate(s.split(',')):
        f = [p.strip() for p in part.split(':')]
        if len(f) != 3:
            raise ValueError(f'--layers entry {i + 1} "{part}": expected Name:thickness_um:n')
        name, t, n = f
        try:
            t_um = float(t)
        except ValueError:
            raise ValueError(f'--layers entry {i + 1} "{part}": thickness "{t}" is not a number')
        if t_um <= 0:
            continue
        try:
            nc, mat = _parse_n(n, lam_nm)
        except Exception:
            raise ValueError(f'--layers entry {i + 1} "{part}": n "{n}" is neither a number '
                             f'nor a known material')
        out.append(dict(name=name or f'Layer {i + 1}', t_um=t_um, n=nc, kappa=nc.imag, mat=mat))
    return out


def legacy_stack(t_lip_um, n_lip, t_aq_um, n_aq, t_mu_um, n_mu, names=None):
    names = list(names) if names else list(LEGACY_NAMES)
    out = []
    for nm, t, n, mat in zip(names, (t_lip_um, t_aq_um, t_mu_um), (n_lip, n_aq, n_mu),
                             ('lipid', 'aqueous', 'mucin')):
        if t is None or float(t) <= 1e-9:
            continue
        nc = complex(n) if n is not None else complex(1.333, 0.0)
        out.append(dict(name=str(nm), t_um=float(t), n=nc, kappa=nc.imag, mat=mat))
    return out


def stack_at_lambda(stack, lam_nm):
    """The stack at another wavelength (reference planar_beads/fdtd2d_core.py:83-95,
    planar_3d/fdtd3d_core.py:92-104): a layer that names a material takes that material's
    real index at ``lam_nm`` and keeps its absorption; layers given by a number stay."""
    if stack is None:
        return None
    from lens_sensor_amd.common.physics import n_at
    out = []
    for layer in stack:
        layer = dict(layer)
        if layer.get('mat'):
            n = complex(n_at(layer['mat'], lam_nm), complex(layer['n']).imag)
            layer['n'], layer['kappa'] = n, n.imag
        out.append(layer)
    return out


def stack_total_um(stack):
    return float(sum(L['t_um'] for L in stack or []))


def stack_bounds(stack, y_top=0.0):
    out, y = [], float(y_top)
    for L in stack or []:
        out.append((L['name'], y, y - L['t_um'], complex(L['n'])))
        y -= L['t_um']
    return out


def legacy_layers_tuple(stack, n_default=1.333):
    st = list(stack or [])
    t = [st[i]['t_um'] * 1e-6 if i < len(st) else 0.0 for i in range(3)]
    n2 = complex(st[1]['n']).real if len(st) > 1 else (complex(st[0]['n']).real if st else n_default)
    return (t[0], t[1], t[2], n2)


def describe(stack):
    if not stack:
        return 'no layers (core directly on the substrate)'
    parts = []
    for L in stack:
        n = complex(L['n'])
        ns = f'{n.real:.4f}' + (f'+{n.imag:.2e}i' if n.imag else '')
        parts.append(f'{L["name"]} {L["t_um"]:g} um (n={ns})')
    return ' | '.join(parts)


def stack_to_npz(stack):
    st = list(stack or [])
    return dict(stack_names=np.array([L['name'] for L in st] or [''], dtype=str)[:len(st)],
                stack_t_um=np.array([L['t_um'] for L in st], dtype=np.float64),
                stack_n_re=np.array([complex(L['n']).real for L in st], dtype=np.float64),
                stack_n_im=np.array([complex(L['n']).imag for L in st], dtype=np.float64))
