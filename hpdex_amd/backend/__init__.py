"""hpdex backend surface on MI355X: mannwhitneyu and group_mean
(reference: src/hpdex/backend/__init__.py)."""
from .mannwhitneyu_ import group_mean, mannwhitneyu

__all__ = ["mannwhitneyu", "group_mean"]
