"""The partial rebuild (`build --only`) refuses to link objects older than a shared header:
every unit includes the kernel-argument structs, so such an object may disagree with the
rebuilt units on their layout."""

import os

import pytest

from jathenax_amd import build


def test_only_refuses_objects_older_than_a_header(tmp_path, monkeypatch):
    pkg = tmp_path / "pkg"
    csrc, obj, inc = pkg / "csrc", pkg / "_build", tmp_path / "include"
    for d in (csrc, obj, inc):
        d.mkdir(parents=True)
    for stem in ("a", "b"):
        (csrc / f"{stem}.hip").write_text("")
        (obj / f"{stem}.o").write_text("")
    header = csrc / "shared.hpp"
    header.write_text("")
    (inc / "api.h").write_text("")
    old = header.stat().st_mtime - 100
    for stem in ("a", "b"):
        os.utime(obj / f"{stem}.o", (old, old))
    monkeypatch.setattr(build, "PKG", pkg)
    monkeypatch.setattr(build, "CSRC", csrc)
    monkeypatch.setattr(build, "OBJ", obj)
    monkeypatch.setattr(build, "_hipcc", lambda: "hipcc")
    with pytest.raises(RuntimeError, match="b.o is older than a shared header"):
        build.build(only=["a"])


def test_only_needs_every_other_object(tmp_path, monkeypatch):
    pkg = tmp_path / "pkg"
    csrc, obj = pkg / "csrc", pkg / "_build"
    csrc.mkdir(parents=True)
    obj.mkdir()
    for stem in ("a", "b"):
        (csrc / f"{stem}.hip").write_text("")
    monkeypatch.setattr(build, "PKG", pkg)
    monkeypatch.setattr(build, "CSRC", csrc)
    monkeypatch.setattr(build, "OBJ", obj)
    monkeypatch.setattr(build, "_hipcc", lambda: "hipcc")
    with pytest.raises(RuntimeError, match="no earlier object for b.o"):
        build.build(only=["a"])
