"""Locations of the in-tree native libraries."""

import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "lib")
ENGINE_LIB = os.path.join(LIB_DIR, "libsimcoproj.so")
GEN_LIB = os.path.join(LIB_DIR, "libsimcogen.so")
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
