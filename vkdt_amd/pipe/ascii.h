/*
 * ascii.h -- ':'-separated field readers of the cfg/params/connectors text
 * formats (src/pipe/asciiio.h): a token is up to 8 characters before the
 * next ':' (it may contain spaces, e.g. "noise a"); numbers are read with
 * strtol/strtof and the separator after them is consumed.
 */
#pragma once
#include "api.h"
#include <stdlib.h>

static inline dt_token_t dt_read_token(char *line, char **out)
{
  char buf[8] = {0};
  int i = 0;
  while(line[i] && line[i] != ':' && line[i] != '\n' && i < 8) { buf[i] = line[i]; i++; }
  dt_token_t t = 0;
  for(int k = 0; k < 8; k++) t |= (dt_token_t)(uint8_t)buf[k] << (8 * k);
  *out = line + i;
  if((*out)[0] == ':' || (*out)[0] == '\n') (*out)++;
  return t;
}

static inline int dt_read_int(char *line, char **out)
{
  if(line[0] == 0) { *out = line; return 0; }
  const int res = (int)strtol(line, out, 10);
  if(**out && *out != line) *out = *out + 1;
  return res;
}

static inline float dt_read_float(char *line, char **out)
{
  if(line[0] == 0) { *out = line; return 0.0f; }
  const float res = strtof(line, out);
  if(**out && *out != line) *out = *out + 1;
  return res;
}
