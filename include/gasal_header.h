/*
 * gasal_header.h -- umbrella header, drop-in for nahmedraja/GASAL2
 * src/gasal_header.h:1-13.
 */
#ifndef __GASAL_HEADER_H__
#define __GASAL_HEADER_H__

#include "gasal.h"
#include "args_parser.h"
#include "gasal_align.h"
#include "host_batch.h"
#include "ctors.h"
#include "interfaces.h"
#include "gasal2_ext.h"
#include "gasal2_stream.h"

#endif
