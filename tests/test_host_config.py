"""Host side of the whole-likelihood path (no GPU): the model description is
checked and turned into the C configuration block before any device call."""
import pytest

from xpsi_amd import configs
from xpsi_amd._lib import XHError
from xpsi_amd.pipeline import model_config


def _cfg(**kw):
    return model_config(dict(configs.config1(), **kw))[0]


def test_defaults_fp64_and_fp32_response_fold():
    cfg = _cfg()
    assert cfg.precision == 0
    assert cfg.fold_kernel == 0
    assert cfg.fold_terms == 0
    assert _cfg(precision="mixed").precision == 1


@pytest.mark.parametrize("terms", [0, 3, 6])
def test_fold_terms_accepted(terms):
    assert _cfg(precision="mixed", fold_terms=terms).fold_terms == terms


@pytest.mark.parametrize("bad", [dict(fold_terms=5), dict(precision="bf16"), dict(fold_kernel="gemm")])
def test_invalid_options_raise(bad):
    with pytest.raises(XHError):
        _cfg(**bad)


def test_hot_regions_must_share_leaves():
    m = configs.config1()
    h = m["hot_regions"][0]
    m["hot_regions"] = [h, dict(h, num_leaves=32)]
    with pytest.raises(XHError):
        model_config(m)
