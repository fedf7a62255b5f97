"""A JSON Schema (draft 2020-12) validator for the keyword subset FoilBench's interchange
schemas use (foilbench_amd/schemas/*.json, the reference's spec/schemas). The reference
validates with the `jsonschema` package (fb/core/_schema_adapter.py); this one reports the
first violation as ``SchemaError(path, message)``.
"""

from __future__ import annotations

import json
import math
import re
from functools import lru_cache
from pathlib import Path

SCHEMA_DIR = Path(__file__).resolve().parent.parent / "schemas"


class SchemaError(ValueError):
    def __init__(self, path: str, message: str) -> None:
        super().__init__(f"{path or '$'}: {message}")
        self.path = path
        self.message = message


@lru_cache(maxsize=None)
def load_schema(name: str) -> dict:
    return json.loads((SCHEMA_DIR / name).read_text(encoding="utf-8"))


def _is_type(value, t: str) -> bool:
    if t == "object":
        return isinstance(value, dict)
    if t == "array":
        return isinstance(value, list)
    if t == "string":
        return isinstance(value, str)
    if t == "boolean":
        return isinstance(value, bool)
    if t == "null":
        return value is None
    if t == "integer":
        return (isinstance(value, int) and not isinstance(value, bool)) or (
            isinstance(value, float) and math.isfinite(value) and value.is_integer())
    if t == "number":
        return isinstance(value, (int, float)) and not isinstance(value, bool)
    raise SchemaError("", f"unsupported type keyword {t!r}")


def _equal(a, b) -> bool:
    if isinstance(a, bool) or isinstance(b, bool):
        return type(a) is type(b) and a == b
    return a == b


def _resolve(root: dict, ref: str) -> dict:
    if not ref.startswith("#/"):
        raise SchemaError("", f"only local references are supported, got {ref!r}")
    node = root
    for part in ref[2:].split("/"):
        node = node[part.replace("~1", "/").replace("~0", "~")]
    return node


def validate(value, schema, root: dict | None = None, path: str = "") -> None:
    """Raise SchemaError at the first violation of `schema` by `value`."""
    root = schema if root is None else root
    if schema is True:
        return
    if schema is False:
        raise SchemaError(path, "no value is allowed here")
    if "$ref" in schema:
        validate(value, _resolve(root, schema["$ref"]), root, path)
    if "type" in schema:
        types = schema["type"] if isinstance(schema["type"], list) else [schema["type"]]
        if not any(_is_type(value, t) for t in types):
            raise SchemaError(path, f"expected {' or '.join(types)}, got {type(value).__name__}")
    if "const" in schema and not _equal(value, schema["const"]):
        raise SchemaError(path, f"must equal {schema['const']!r}")
    if "enum" in schema and not any(_equal(value, e) for e in schema["enum"]):
        raise SchemaError(path, f"must be one of {schema['enum']!r}")
    if isinstance(value, (int, float)) and not isinstance(value, bool):
        if "minimum" in schema and value < schema["minimum"]:
            raise SchemaError(path, f"must be >= {schema['minimum']}")
        if "maximum" in schema and value > schema["maximum"]:
            raise SchemaError(path, f"must be <= {schema['maximum']}")
        if "exclusiveMinimum" in schema and value <= schema["exclusiveMinimum"]:
            raise SchemaError(path, f"must be > {schema['exclusiveMinimum']}")
        if "exclusiveMaximum" in schema and value >= schema["exclusiveMaximum"]:
            raise SchemaError(path, f"must be < {schema['exclusiveMaximum']}")
    if isinstance(value, str):
        if "minLength" in schema and len(value) < schema["minLength"]:
            raise SchemaError(path, f"must have at least {schema['minLength']} characters")
        if "pattern" in schema and re.search(schema["pattern"], value) is None:
            raise SchemaError(path, f"must match {schema['pattern']!r}")
    if isinstance(value, list):
        if "minItems" in schema and len(value) < schema["minItems"]:
            raise SchemaError(path, f"must have at least {schema['minItems']} items")
        if "maxItems" in schema and len(value) > schema["maxItems"]:
            raise SchemaError(path, f"must have at most {schema['maxItems']} items")
        if schema.get("uniqueItems"):
            for i in range(len(value)):
                for j in range(i):
                    if _equal(value[i], value[j]):
                        raise SchemaError(path, "items must be unique")
        prefix = schema.get("prefixItems", [])
        for k, (item, sub) in enumerate(zip(value, prefix)):
            validate(item, sub, root, f"{path}[{k}]")
        if "items" in schema:
            for k in range(len(prefix), len(value)):
                validate(value[k], schema["items"], root, f"{path}[{k}]")
    if isinstance(value, dict):
        for key in schema.get("required", []):
            if key not in value:
                raise SchemaError(path, f"missing required property {key!r}")
        props = schema.get("properties", {})
        for key, item in value.items():
            if key in props:
                validate(item, props[key], root, f"{path}.{key}")
            elif "additionalProperties" in schema:
                validate(item, schema["additionalProperties"], root, f"{path}.{key}")
    for sub in schema.get("allOf", []):
        validate(value, sub, root, path)
    if "anyOf" in schema:
        if not any(_ok(value, sub, root, path) for sub in schema["anyOf"]):
            raise SchemaError(path, "matches none of the anyOf alternatives")
    if "oneOf" in schema:
        n = sum(_ok(value, sub, root, path) for sub in schema["oneOf"])
        if n != 1:
            raise SchemaError(path, f"must match exactly one oneOf alternative (matched {n})")
    if "if" in schema:
        branch = "then" if _ok(value, schema["if"], root, path) else "else"
        if branch in schema:
            validate(value, schema[branch], root, path)


def _ok(value, schema, root, path) -> bool:
    try:
        validate(value, schema, root, path)
    except SchemaError:
        return False
    return True
