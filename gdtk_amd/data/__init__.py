"""Gas-model and reaction-mechanism data (JSON) for the chemistry path."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name: str) -> dict:
    with open(os.path.join(HERE, name if name.endswith(".json") else name + ".json")) as f:
        return json.load(f)
