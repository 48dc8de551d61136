"""The library's deck reader (host only, through the C-ABI qz_deck_keys): the reference's input format
(src/Config.hpp:1-11, README.md section 5) on the forms its decks use -- flat keys, nested blocks, named blocks
that join a list, ';' and comments, quotes, units with the reference density -- and the errors it must raise."""
import math

import pytest


@pytest.fixture(scope="module")
def qa():
    import quarz_amd
    quarz_amd.lib()
    return quarz_amd


def test_flat_blocks_and_named_blocks(qa):
    deck = """
    xi.step = 0.01                      # flat
    xi { min = -12; max = 0 }           # a block prefixes its keys
    grid {
      type = regions
      sub { a = 1 }                     # nested
    }
    bunch driver {                      # joins beams
      n0 = 2;  sigma_r = 0.3
    }
    beam witness { n0 = 10 }
    species electrons { ppc = 4 }       # joins plasma.species
    diag light { every = 10 }           # keys under diag.light.
    """
    k = qa.deck_keys(deck)
    assert k["xi.step"] == "0.01" and k["xi.min"] == "-12" and k["xi.max"] == "0"
    assert k["grid.type"] == "regions" and k["grid.sub.a"] == "1"
    assert k["driver.n0"] == "2" and k["driver.sigma_r"] == "0.3" and k["witness.n0"] == "10"
    assert k["beams"].split() == ["driver", "witness"]
    assert k["electrons.ppc"] == "4" and k["plasma.species"] == "electrons"
    assert k["diag.light.every"] == "10" and k["diag.names"] == "light"


def test_list_join_keeps_a_given_list(qa):
    k = qa.deck_keys("beams = witness\nbunch driver { n0 = 1 }\nbunch witness { n0 = 2 }\n")
    assert k["beams"].split() == ["witness", "driver"]


def test_quotes_comments_and_overrides(qa):
    deck = 'a = "x; y"   # the ; inside quotes does not end the statement\nb = 1 # c = 2\nc = 3\n'
    k = qa.deck_keys(deck, {"c": 4, "d": "five"})
    assert k["a"] == '"x; y"' and k["b"] == "1" and "c = 2" not in k
    assert k["c"] == "4" and k["d"] == "five"     # an override replaces, and may add


@pytest.mark.parametrize("deck,words", [
    ("a = 1\na = 2\n", "second time"),
    ("x { a = 1\n", "no closing"),
    ("a = 1 }\n", "closes no block"),
    ("just words\n", "key = value"),
    ("one two three { a = 1 }\n", "block header"),
    ("= 3\n", "without a key"),
    ("x { y { a = 1 } }\nx.y.a = 2\n", "second time"),     # a block key and its flat form are one key
])
def test_errors(qa, deck, words):
    with pytest.raises(qa.QuarzError) as e:
        qa.deck_keys(deck)
    assert words in str(e.value)
