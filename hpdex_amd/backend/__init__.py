"""hpdex backend surface on MI355X: mannwhitneyu, group_mean and the numba
pipeline's driver (reference: src/hpdex/backend/__init__.py)."""
from .de_ import parallel_differential_expression as parallel_differential_expression_numba
from .mannwhitneyu_ import group_mean, mannwhitneyu

__all__ = ["mannwhitneyu", "group_mean", "parallel_differential_expression_numba"]
