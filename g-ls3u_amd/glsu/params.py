"""WFF parameter helpers (argument order of WFT/WFT2_CPU.h:30-36: sigma, low, high, step)."""
from ._lib import WffParams, check, lib


def wff_params(sigma_x=10.0, wxl=-0.2, wxh=0.2, wxi=0.1, sigma_y=None, wyl=None, wyh=None,
               wyi=None, thr=-1.0):
    """WFF parameters; the y axis defaults to the x axis values.  thr < 0 selects
    6*sqrt(mean|f|^2/3) computed on the first call and cached
    (WFT/WFT2_CPU.cpp:977-1022)."""
    return WffParams(float(sigma_x), float(wxl), float(wxh), float(wxi),
                     float(sigma_x if sigma_y is None else sigma_y),
                     float(wxl if wyl is None else wyl),
                     float(wxh if wyh is None else wyh),
                     float(wxi if wyi is None else wyi), float(thr))


def dpra_reference_params():
    """The WFF parameters DPRA_CPU / DPRA_CUDAF hard-code (DPRA/dpra_cpu.cpp:23-25)."""
    p = WffParams()
    check(lib().glsu_wff_params_dpra_reference(p))
    return p


def wft2_default_params():
    """Defaults of WFT2_cpu(W, H, WFF, z, n) (WFT/WFT2_CPU.cpp:43-54)."""
    p = WffParams()
    check(lib().glsu_wff_params_default(p))
    return p
