#!/bin/sh
# Writes scene_cache_keys.txt: for each shipped scene under scenes/, the key of the reference's matrix
# cache and the JSON it hashes, as printed by oracle/_ref/config_key (built against the reference's own
# vendored JSON library: make -C oracle _ref/config_key).
cd "$(dirname "$0")" || exit 1
: > scene_cache_keys.txt
for f in scenes/*.json; do
    ../../oracle/_ref/config_key "$f" | sed "s|^|$(basename "$f")\t|" >> scene_cache_keys.txt || exit 1
done
