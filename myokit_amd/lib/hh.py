"""
Hodgkin-Huxley gate detection for Rush-Larsen updates.

Follows ``myokit/lib/hh.py``: ``convert_hh_states_to_inf_tau_form:922-985``
(adds nested ``tau = 1 / (alpha + beta)`` and ``inf = alpha * tau`` and
rewrites ``dot(x) = (inf - x) / tau``), ``get_alpha_and_beta:988-1096``
(shape ``alpha * (1 - x) - beta * x`` with either operand order, alpha and
beta plain variable references that depend on no state other than ``v``) and
``get_inf_and_tau:1099-1191`` (shape ``(inf - x) / tau``).
"""
from .. import _expressions as ex
from .._model import Model, Variable


def _membrane_potential(model, v):
    if v is None:
        v = model.label('membrane_potential')
        if v is None:
            raise ValueError(
                'Membrane potential must be given as `v` or by setting the'
                ' label `membrane_potential` in the model.')
        return v
    if isinstance(v, Variable):
        v = v.qname()
    return model.get(v)


def _depends_only_on_v(term, model, v):
    rhs = term.rhs().clone(expand=True, retain=(ex.Name(v),))
    for state in model.states():
        if state is not v and rhs.depends_on(ex.Name(state)):
            return False
    return True


def get_alpha_and_beta(x, v=None):
    """Returns ``(alpha, beta)`` variables if ``dot(x)`` has the shape
    ``alpha * (1 - x) - beta * x``, else ``None``."""
    if not x.is_state():
        return None
    rhs = x.rhs()
    if not isinstance(rhs, ex.Minus):
        return None
    a, b = rhs
    if not (isinstance(a, ex.Multiply) and isinstance(b, ex.Multiply)):
        return None
    xname = ex.Name(x)
    one_minus_x = ex.Minus(ex.Number(1), xname)
    if a[0] == one_minus_x:
        alpha = a[1]
    elif a[1] == one_minus_x:
        alpha = a[0]
    else:
        return None
    if b[0] == xname:
        beta = b[1]
    elif b[1] == xname:
        beta = b[0]
    else:
        return None
    if not (isinstance(alpha, ex.Name) and isinstance(beta, ex.Name)):
        return None
    alpha, beta = alpha.var(), beta.var()
    if alpha.is_state() or beta.is_state():
        return None
    model = x.model()
    v = _membrane_potential(model, v)
    for term in (alpha, beta):
        if not _depends_only_on_v(term, model, v):
            return None
    return alpha, beta


def get_inf_and_tau(x, v=None):
    """Returns ``(inf, tau)`` variables if ``dot(x)`` has the shape
    ``(inf - x) / tau``, else ``None``."""
    if not x.is_state():
        return None
    rhs = x.rhs()
    if not isinstance(rhs, ex.Divide):
        return None
    numer, tau = rhs
    if not isinstance(tau, ex.Name) or not isinstance(numer, ex.Minus):
        return None
    inf, xname = numer
    if not (isinstance(inf, ex.Name) and isinstance(xname, ex.Name)):
        return None
    if xname != ex.Name(x):
        return None
    inf, tau = inf.var(), tau.var()
    if inf.is_state() or tau.is_state():
        return None
    model = x.model()
    v = _membrane_potential(model, v)
    for term in (inf, tau):
        if not _depends_only_on_v(term, model, v):
            return None
    return inf, tau


def convert_hh_states_to_inf_tau_form(model, v=None):
    """Returns a copy of ``model`` with every alpha/beta gate rewritten to
    the inf/tau form."""
    if not isinstance(model, Model):
        raise ValueError('Given `model` must be a myokit.Model.')
    model = model.clone()
    v = _membrane_potential(model, v)
    for x in list(model.states()):
        res = get_alpha_and_beta(x, v)
        if res is None:
            continue
        a, b = ex.Name(res[0]), ex.Name(res[1])
        tau = x.add_variable_allow_renaming('tau')
        tau.set_rhs(ex.Divide(ex.Number(1), ex.Plus(a, b)))
        inf = x.add_variable_allow_renaming('inf')
        inf.set_rhs(ex.Multiply(a, ex.Name(tau)))
        x.set_rhs(ex.Divide(ex.Minus(ex.Name(inf), ex.Name(x)),
                            ex.Name(tau)))
    model.validate()
    return model


def get_rl_expression(x, dt, v=None):
    """``inf + (x - inf) * exp(-(dt / tau))`` for an inf/tau state."""
    if not isinstance(dt, ex.Expression):
        raise ValueError('Argument `dt` must be a myokit.Expression.')
    res = get_inf_and_tau(x, v)
    if res is None:
        return None
    xn, inf, tau = ex.Name(x), ex.Name(res[0]), ex.Name(res[1])
    return ex.Plus(inf, ex.Multiply(
        ex.Minus(xn, inf), ex.Exp(ex.PrefixMinus(ex.Divide(dt, tau)))))
