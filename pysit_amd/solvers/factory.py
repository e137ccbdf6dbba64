"""Solver registration and selection.

Mirrors pysit/util/basic_registration_factory.py:19-196,
pysit/util/solvers.py:6-45 and pysit/solvers/solver_factory.py:10-105:
solver classes declare a ``supports`` dict (merged down the class hierarchy by
``inherit_dict``); the factory passes the call's keyword arguments (plus its
own defaults) to every registered class's ``_factory_validation_function``,
which answers DoesNotMatch / IncompleteMatch / CompleteMatch.  One complete
match wins; otherwise the first incomplete match in registration order;
none raises ``NoMatchError``; several complete matches raise
``MultipleMatchError``.  Unknown keyword arguments are ignored.
"""
from collections import OrderedDict

__all__ = ["supports", "inherit_dict", "NoMatchError", "MultipleMatchError", "ValidationFunctionError",
           "DoesNotMatch", "IncompleteMatch", "CompleteMatch", "BasicRegistrationFactory",
           "TernaryRegistrationFactory", "SolverFactory", "PositiveEvenIntegers"]

DoesNotMatch = 0
IncompleteMatch = 1
CompleteMatch = 2


class NoMatchError(Exception):
    pass


class MultipleMatchError(Exception):
    pass


class ValidationFunctionError(AttributeError):
    pass


class _PositiveEvenIntegers(object):
    def __contains__(self, needle):
        return needle > 0 and not needle % 2


PositiveEvenIntegers = _PositiveEvenIntegers()


def supports(needle, haystack):
    """Is ``needle`` (a string or an int) one of the supported values?"""
    if isinstance(needle, str):
        return needle == haystack if isinstance(haystack, str) else needle in haystack
    if isinstance(needle, int):
        return needle == haystack if isinstance(haystack, int) else needle in haystack
    raise ValueError("Needle does not match known requirement type.")


def inherit_dict(name, local_name=None):
    """Class decorator: ``cls.<name>`` = merge of the bases' dicts, then the
    class's own ``<local_name>`` (default ``_<name>``)."""
    local_name = local_name or "_" + name

    def wrap(cls):
        merged = OrderedDict()
        for base in cls.__bases__:
            merged.update(getattr(base, name, dict()))
        merged.update(getattr(cls, local_name, dict()))
        setattr(cls, name, merged)
        return cls

    return wrap


class BasicRegistrationFactory(object):
    def __init__(self, default_widget_type=None, additional_validation_functions=[]):
        self.registry = OrderedDict()
        self.default_widget_type = default_widget_type
        self.validation_functions = ["_factory_validation_function"] + list(additional_validation_functions)

    def __call__(self, *args, **kwargs):
        return self._check_registered_widget(*args, **kwargs)(*args, **kwargs)

    def _check_registered_widget(self, *args, **kwargs):
        hits = [k for k in self.registry if self.registry[k](*args, **kwargs)]
        if not hits:
            if self.default_widget_type is None:
                raise NoMatchError("No types match specified arguments and no default is set.")
            return self.default_widget_type
        if len(hits) > 1:
            raise MultipleMatchError(f"Too many candidate types identified ({len(hits)}).")
        return hits[0]

    def register(self, WidgetType, validation_function=None, is_default=False):
        if is_default:
            self.default_widget_type = WidgetType
            return
        if validation_function is not None:
            if not callable(validation_function):
                raise AttributeError("Keyword argument 'validation_function' must be callable.")
            self.registry[WidgetType] = validation_function
            return
        for name in self.validation_functions:
            if hasattr(WidgetType, name):
                f = getattr(WidgetType, name)
                if getattr(f, "__self__", None) is not WidgetType:
                    raise ValidationFunctionError(f"{WidgetType.__name__}.{name} must be a classmethod.")
                self.registry[WidgetType] = f
                return
        raise ValidationFunctionError(f"No proper validation function for class {WidgetType.__name__} found.")

    def unregister(self, WidgetType):
        self.registry.pop(WidgetType)


class TernaryRegistrationFactory(BasicRegistrationFactory):
    def __init__(self, property_dict_name="supports", **kwargs):
        BasicRegistrationFactory.__init__(self, **kwargs)
        self.property_dict_name = property_dict_name

    def _check_registered_widget(self, *args, **kwargs):
        complete, incomplete = [], []
        for key in self.registry:
            m = self.registry[key](*args, **kwargs)
            if m == CompleteMatch:
                complete.append(key)
            elif m == IncompleteMatch:
                incomplete.append(key)
        if len(complete) == 1:
            return complete[0]
        if len(complete) > 1:
            raise MultipleMatchError(f"Too many candidate types match ({len(complete)}).")
        if incomplete:
            return incomplete[0]
        if self.default_widget_type is None:
            raise NoMatchError("No types match specified arguments and no default is set.")
        return self.default_widget_type


class SolverFactory(TernaryRegistrationFactory):
    """Factory whose own ``supports`` defaults are merged under the caller's kwargs."""

    supports = {"equation_physics": None, "equation_dynamics": None}

    def _check_registered_widget(self, *args, **kwargs):
        merged = dict(self.supports)
        merged.update(kwargs)
        return TernaryRegistrationFactory._check_registered_widget(self, *args, **merged)

    def register(self, WidgetType, validation_function=None, is_default=False):
        for key in ("equation_physics", "equation_dynamics"):
            if self.supports[key] != WidgetType.supports[key]:
                raise ValueError(f"Factory {type(self).__name__} accepts {key} '{self.supports[key]}' "
                                 f"but {WidgetType.__name__} supports '{WidgetType.supports[key]}'.")
        TernaryRegistrationFactory.register(self, WidgetType, validation_function, is_default)
