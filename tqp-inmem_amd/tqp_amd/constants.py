"""Dtype policy and sentinels of the relational hot path.

Mirrors constants.py:3-8 of the reference: strings are int8 byte matrices, integers int64, floats
float64, dates int32 day counts since 1990-01-01; `eps` widens float `<=`/`>=` literals and is the
`atol` of float `=`; `null` (-64) is the NULL sentinel that `count` skips and right-outer joins emit.
"""
import torch

string_dtype = torch.int8
int_dtype = torch.int64
float_dtype = torch.float64
date_dtype = torch.int32

eps = 1e-5
null = -64
